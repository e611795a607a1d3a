"""The reference's own printed ordinate tables pin the quadrature generator.

Fixtures: tests/golden/ordinates.json, extracted by tests/golden/make_ordinates.py
from the lb.out logs the reference ships (16 printed decimals).
"""
import json
import math
import os

import numpy as np

from lionbolt_amd import angular

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ordinates.json")


def _load():
    with open(GOLD) as fh:
        return json.load(fh)


def test_snprep_matches_reference_tables():
    d = _load()
    for name in ("L7_3d", "L11_3d"):
        ref = np.array(d[name]["k"])
        s = angular.snprep(d[name]["L"])
        assert ref.shape == s.k.shape
        assert np.abs(s.k - ref).max() <= 6e-16, name


def test_gauss_legendre_matches_slab_table():
    d = _load()["L15_slab"]
    mu, w = angular.gauss_legendre(d["L"] + 1)
    assert np.abs(mu - np.array(d["k"])[:, 0]).max() <= 6e-16
    assert abs(w.sum() - 2.0) <= 1e-15


def test_quadrature_moments():
    for ang in (angular.snprep(3), angular.product(8, 32, 3), angular.level_symmetric_s8()):
        assert abs(ang.w.sum() - 4 * math.pi) <= 1e-13
        assert np.allclose(np.linalg.norm(ang.k, axis=1), 1.0, atol=1e-15)
        for c in range(3):
            assert abs(np.sum(ang.w * ang.k[:, c] ** 2) - 4 * math.pi / 3) <= 1e-13
            assert abs(np.sum(ang.w * ang.k[:, c])) <= 1e-13
    s8 = angular.level_symmetric_s8()
    assert s8.NI == 80
    for c in range(3):   # LQ_n S8 integrates mu^4 exactly: 4 pi / 5
        assert abs(np.sum(s8.w * s8.k[:, c] ** 4) - 4 * math.pi / 5) <= 1e-13
