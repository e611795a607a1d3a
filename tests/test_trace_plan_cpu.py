"""The LDS slot plan of the direction-resident LDS sweep, checked on the CPU by
replaying what the kernel does with it, level by level, on an array that
stands for the workgroup's LDS and records which element's flux each slot holds:

* every upstream trace the plan sends to LDS finds its producer's flux there;
* two cells of one level never write the same slot, and no slot is both read
  and written in one level (its waves run unordered); cells write ring slots
  and staging copies write staging slots of the consumer level's parity only;
* a staged producer lies two or more levels before its consumer (its global
  store is ordered by an earlier barrier);
* the traces are counted once each in the plan's statistics;
* with the FP64 flux stored only for the cells the plan marks (the FP32 flux
  mode), every read from global memory finds its producer stored: the staging
  copies, and all traces of a cell with any trace outside LDS.

The upstream faces and levels come from the normals and schedule of
lb_schedule_host, recomputed here with the reference's rule (k.n < -EPSILON).
Small LDS sizes force evictions, staging and global reads on small meshes."""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M
from lionbolt_amd.ops import schedule_host, trace_plan_host

EPS = np.finfo(np.float64).eps


def replay(m, ang, G, lds_bytes):
    sl, nlev, nrm = schedule_host(m, ang)
    p = trace_plan_host(m, ang, G, lds_bytes)
    ring, S = p["ring"], p["stage_slots"]
    assert (ring + 2 * S) * 32 * G <= lds_bytes
    nbr = np.asarray(m.nbr)
    total = 0
    for i in range(ang.NI):
        k = ang.k[i]
        kn = (k[0] * nrm[:, :, 0] + k[1] * nrm[:, :, 1]) + k[2] * nrm[:, :, 2]
        up = (nbr >= 0) & (kn < -EPS)
        order = sl[i]
        lev = np.zeros(m.NE, np.int64)
        for e in order:
            u = nbr[e][up[e]]
            lev[e] = lev[u].max() + 1 if len(u) else 0
        lv = lev[order]
        assert np.all(np.diff(lv) >= 0) and lv[-1] == nlev[i] - 1
        off = np.searchsorted(lv, np.arange(nlev[i] + 1))
        us, ws, so = p["uslot"][i], p["wslot"][i], p["stage_off"][i]
        lds = np.full(ring + 2 * S, -1, np.int64)
        stored = np.zeros(m.NE, bool)   # FP64 flux in global memory, FP32 flux mode
        gp = p["gprod"][i]
        for l in range(nlev[i]):
            pos = np.arange(off[l], off[l + 1])
            read = set()
            for s in pos:
                e = order[s]
                if np.any(us[s][up[e]] < 0):   # read from global memory, every upstream trace
                    assert np.all(stored[nbr[e][up[e]]]), (i, l, e)
                for f in range(4):
                    v = us[s, f]
                    if not up[e, f]:
                        assert v == -1
                        continue
                    q = nbr[e, f]
                    if v < 0:
                        continue
                    assert lds[v] == q, (i, l, e, f, v)
                    read.add(int(v))
                    if v >= ring:
                        assert ring + (l & 1) * S <= v < ring + (l & 1) * S + S
                        assert lev[q] <= l - 2
            w = ws[pos]
            w = w[w >= 0]
            assert np.all(w < ring) and len(np.unique(w)) == len(w)
            lds[w] = order[pos][ws[pos] >= 0]
            stored[order[pos][gp[pos] != 0]] = True
            ent = p["stage"][so[l + 1]:so[l + 2]]          # copies made during level l for level l + 1
            # waves of one level run unordered: no slot is both read and written in it
            assert not read & (set(w.tolist()) | set(ent[:, 0].tolist()))
            assert len(ent) <= S
            if len(ent):
                assert np.all((ent[:, 0] >= ring + ((l + 1) & 1) * S) & (ent[:, 0] < ring + ((l + 1) & 1) * S + S))
                assert len(np.unique(ent[:, 0])) == len(ent)
                assert np.all(lev[ent[:, 1]] <= l - 1)
                assert np.all(stored[ent[:, 1]])
                lds[ent[:, 0]] = ent[:, 1]
        total += int(up.sum())
    assert p["stats"].sum() == total
    return p


@pytest.fixture(scope="module")
def small():
    return M.jittered_cube(5), angular.snprep(3)


def test_ample_lds_reads_every_trace_from_the_ring(small):
    m, ang = small
    p = replay(m, ang, 1, 163840)
    assert p["stats"][0] == p["stats"].sum()
    assert not p["gprod"].any()


@pytest.mark.parametrize("G,slots", [(1, 64), (1, 96), (1, 128), (2, 100), (4, 48)])
def test_tight_lds_plan_is_consistent(small, G, slots):
    m, ang = small
    p = replay(m, ang, G, slots * 32 * G)
    ringr, staged, late, over = p["stats"]
    assert staged > 0 and ringr > 0
    assert 0 < p["gprod"].mean() < 1
    assert ringr + staged > 0.5 * p["stats"].sum()


def test_plan_rejects_an_lds_too_small_for_one_trace(small):
    from lionbolt_amd._lib import LionboltError
    m, ang = small
    with pytest.raises(LionboltError):
        trace_plan_host(m, ang, 4, 100)
