"""Synthetic workloads: the batch union mesh, the uneven-beam draw and the cube beam's
uncollided flux (CPU only)."""
import math

import numpy as np

from lionbolt_amd import mesh as M, problems


def test_union_neighbours_match_a_fresh_face_search():
    parts = [M.jittered_cube(2, a=1.0, seed=11), M.jittered_cube(3, a=2.0, seed=12), M.jittered_cube(1, a=5.0)]
    um, eoff = M.union(parts)
    assert list(eoff) == [0, 48, 48 + 162, 48 + 162 + 6]
    assert np.array_equal(um.nbr, M.build_neighbors(um.node))
    for p in range(3):
        assert np.all(um.mat[eoff[p]:eoff[p + 1]] == p + 1)
        blk = um.nbr[eoff[p]:eoff[p + 1]]
        inside = blk[blk >= 0]
        assert inside.min() >= eoff[p] and inside.max() < eoff[p + 1]


def test_uneven_beams_draw_is_seeded_and_in_range():
    a = problems.uneven_beams(nprob=64)
    b = problems.uneven_beams(nprob=64)
    assert [p["n"] for p in a] == [p["n"] for p in b]
    for p in a:
        assert 7 <= p["n"] <= 21 and 5.0 <= p["a"] <= 20.0
        assert 0.05 <= p["sigt"] <= 1.0 and 0.5 <= p["c"] <= 0.95
        assert abs(np.linalg.norm(p["k0"]) - 1.0) < 1e-14 and 0.5 <= p["k0"][2] <= 1.0
        assert p["mesh"].NE == 6 * p["n"] ** 3
    ne = np.array([p["mesh"].NE for p in a])
    assert ne.min() >= 6 * 7 ** 3 and ne.max() <= 6 * 21 ** 3 and len(set(ne)) > 8
