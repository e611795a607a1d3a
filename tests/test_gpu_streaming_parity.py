"""The many-group sweep path against the oracle.

With 8 or more groups per cell the sweep kernels read records, face
coefficients and sources with non-temporal loads (the streaming instantiations
of the level kernel, the direction-resident walk and the LDS walk), and the
multigroup solve evaluates psi and its residual 8 groups per sweep.  Here every
group slice of such sweeps is compared with the oracle's single-group sweep
(TransportAlgebra.f90:29-106) to 1e-12 and with the single-group GPU sweep of
that group bit for bit, in every launch mode; the multigroup solve with its
8-group tail and with 8-group coupled batches is compared with the oracle's
group loop (Main.f90:326-476, 584-629)."""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M, problems
from oracle.oracle import OracleOp

pytestmark = pytest.mark.gpu

MODES = [("levels", 0), ("graph", 0), ("direction", 256), ("direction", 1024), ("direction_lds", 512), ("auto", 0)]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("G", [8, 9, 16, 24])
def test_many_group_sweeps_match_oracle(G):
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(4, a=4.0, seed=1001 + G, slab=(1.5, 2.5))
    ang = angular.product(4, 8, 1)
    sigt = np.stack([np.geomspace(0.05, 2.0, G), 3.0 * np.geomspace(0.05, 2.0, G)], axis=1)
    rng = np.random.default_rng(G)
    q = rng.standard_normal((ang.NI, m.NE, 4, G))
    ref = []
    orc = OracleOp(m, ang)
    for g in range(G):
        orc.set_physics(sigt[g], np.zeros((2, ang.L + 1)))
        ref.append(orc.transport_matinv(q[..., g].reshape(-1)).reshape(ang.NI, m.NE, 4))
    op = Operator(m, ang)
    op.set_physics_groups(sigt, np.zeros((G, 2, ang.L + 1)))
    qd = torch.from_numpy(q.reshape(-1).copy()).cuda()
    singles = [torch.from_numpy(np.ascontiguousarray(q[..., g]).reshape(-1)).cuda() for g in range(G)]
    for mode, block in MODES:
        op.set_sweep_mode(mode)
        if block:
            op.set_direction_block(block)
        reps = 10 if mode == "auto" else 1          # the autotuner's timed sweeps and its settled choice
        for _ in range(reps):
            psi = torch.full_like(qd, float("nan"))
            op.transport_matinv_dev(qd, psi, 0, G)
            op.synchronize()
            out = psi.cpu().numpy().reshape(ang.NI, m.NE, 4, G)
            for g in range(G):
                assert rel(out[..., g], ref[g]) <= 1e-12, (mode, block, g)
        for g in (0, G // 2, G - 1):
            one = torch.full_like(singles[g], float("nan"))
            op.transport_matinv_dev(singles[g], one, g, 1)
            op.synchronize()
            assert np.array_equal(one.cpu().numpy().reshape(ang.NI, m.NE, 4), out[..., g]), (mode, block, g)
    op.close()


def beam_qmom(m, ang, S, sigt):
    """The +z beam's group source moments (problems.beam_group_sources) as a host array."""
    from lionbolt_amd.ops import Operator
    op = Operator(m, ang)
    q = problems.beam_group_sources(op, m, S, sigt)[0].cpu().numpy()
    op.close()
    return q


def _oracle_multigroup(m, ang, sigt, S, qmom, rtol):
    G = S.shape[0]
    _, Mm = angular.moment_matrices(ang)
    D, _ = angular.moment_matrices(ang)
    within = OracleOp(m, ang)
    coupling = OracleOp(m, ang)
    psis, phis = [], []
    for g in range(G):
        s = (Mm @ qmom[g]).reshape(-1)
        for gp in range(g):
            coupling.set_physics(np.ones(sigt.shape[1]), S[gp, g])
            s += coupling.scatter_matvec(psis[gp])
        within.set_physics(sigt[g], S[g, g])
        x, st = within.boltzmann_matinv(s, kind="gmres", m=30, rtol=rtol)
        assert st["status"] == 0
        psis.append(x)
        phis.append(D @ x.reshape(ang.NI, -1))
    return phis


@pytest.mark.parametrize("batch", [1, 8])
def test_ten_group_solve_with_eight_group_sweeps_matches_oracle(batch):
    """Ten groups: group by group, psi and the true residuals come from an
    8-group sweep and a 2-group sweep; with coupled batches of 8 every Arnoldi
    step of the first batch is an 8-group sweep."""
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(4, a=10.0, seed=1001, slab=(4.0, 6.0))
    ang = angular.product(2, 8, 1)
    G = 10
    sigt, S = problems.downscatter_xs(G=G, L=ang.L)
    sigt, S = sigt * 10.0, S * 10.0                   # optically thicker so groups couple noticeably
    qmom = beam_qmom(m, ang, S, sigt)
    ref = _oracle_multigroup(m, ang, sigt, S, qmom, rtol=1e-11)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, problems.within_group(S))
    op.set_group_batch(batch)
    phi = torch.zeros((G, ang.n_moments, m.NENK), dtype=torch.float64, device="cuda")
    stats = op.multigroup_solve_dev(S, phi, qmom=torch.from_numpy(qmom).cuda(), rtol=1e-11, atol=0.0)
    assert all(s["status"] == 0 for s in stats)
    assert max(s["true_residual"] for s in stats) <= 1e-10
    got = phi.cpu().numpy()
    for g in range(G):
        assert rel(got[g], ref[g]) <= 1e-7, g
    op.close()
