"""Pins for the oracle's sweep order (SweepOrder.f90:105-207).

The reference's Kahn passes produce "levels concatenated, ascending element
index inside each level", where an element's level is one more than the
deepest upstream neighbour (0 when it has none).  The test computes those
longest-path levels by fixed-point iteration, a different algorithm, and
compares the full sweep list; it also checks the invariants.
"""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M
from oracle.oracle import OracleOp

EPS = np.finfo(float).eps


def brute_levels(m, nrm, k):
    up = (m.nbr >= 0) & (np.einsum("efc,c->ef", nrm, k) < -EPS)
    lev = np.zeros(m.NE, dtype=np.int64)
    for _ in range(m.NE + 1):
        dep = np.where(up, lev[np.maximum(m.nbr, 0)] + 1, 0).max(axis=1)
        if np.array_equal(dep, lev):
            break
        lev = dep
    else:
        raise AssertionError("dependency cycle")
    return lev, up


@pytest.mark.parametrize("ang_name", ["snprep3", "s8"])
def test_sweep_order_matches_longest_path_levels(small_mesh, ang_name):
    ang = angular.snprep(3) if ang_name == "snprep3" else angular.level_symmetric_s8()
    m = small_mesh
    op = OracleOp(m, ang)
    nrm = op.geometry()["nrm"]
    sl, nlev, lev = op.schedule()
    for i in range(ang.NI):
        L, up = brute_levels(m, nrm, ang.k[i])
        expect = np.lexsort((np.arange(m.NE), L))
        assert np.array_equal(sl[i], expect)
        assert nlev[i] == L.max() + 1
        assert np.array_equal(lev[i], L[expect])
        # every upstream neighbour sits at a strictly lower level
        e, f = np.nonzero(up)
        assert np.all(L[m.nbr[e, f]] < L[e])
        # sl is a permutation
        assert np.array_equal(np.sort(sl[i]), np.arange(m.NE))


def test_upstream_faces_are_interior_only(small_mesh):
    ang = angular.snprep(1)
    m = small_mesh
    op = OracleOp(m, ang)
    nrm = op.geometry()["nrm"]
    for i in range(ang.NI):
        _, up = brute_levels(m, nrm, ang.k[i])
        # each interior face is upstream for exactly one side unless grazing
        kn = np.einsum("efc,c->ef", nrm, ang.k[i])
        inter = m.nbr >= 0
        assert np.all(~up[~inter])
        assert int(up.sum()) == int((inter & (kn < -EPS)).sum())
