"""Pins for the oracle's geometry precompute (FiniteElementAnalysis / PostProcess).

For linear tetrahedra the reference's quadrature-built, mass-eliminated inner
products have closed forms: M = V/20 (I + J), so M^-1 = (20/V)(I - J/5);
the convection matrix is (V/4) grad(u_j) 1^T, so IC_d = 5 (d_d u_a) 1^T; the
triangle face mass is A/12 (I + J).  The normals satisfy sum_f A_f n_f = 0,
point away from the element, and neighbours hold exact negations.
"""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M
from oracle.oracle import OracleOp

FN = M.FACE_NODES
OPP = M.FACE_OPPOSITE


@pytest.fixture(scope="module")
def geo(small_mesh):
    op = OracleOp(small_mesh, angular.snprep(1))
    return small_mesh, op.geometry()


def _grads(m):
    x = m.rg[m.node]
    J = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)
    Jinv = np.linalg.inv(J)                    # rows = grad lambda_1..3
    g = np.empty((m.NE, 4, 3))
    g[:, 1:] = Jinv
    g[:, 0] = -Jinv.sum(axis=1)
    return g


def test_detj_is_six_volumes(geo):
    m, g = geo
    assert np.allclose(g["detJ"], 6.0 * m.volumes(), rtol=1e-13, atol=0)


def test_convection_closed_form(geo):
    m, g = geo
    grad = _grads(m)
    expect = 5.0 * np.transpose(grad, (0, 2, 1))[:, :, :, None] * np.ones(4)[None, None, None, :]
    scale = np.abs(expect).max()
    assert np.abs(g["IC"] - expect).max() <= 1e-13 * scale


def test_face_mass_closed_form(geo):
    m, g = geo
    V = m.volumes()
    for f in range(4):
        c = 5.0 * g["area"][:, f] / (3.0 * V)
        E = np.zeros((4, 4))
        for b in FN[f]:
            for a in range(4):
                E[a, b] = int(a == b) + int(a in FN[f]) - 0.8
        expect = c[:, None, None] * E[None]
        assert np.abs(g["IFM"][:, f] - expect).max() <= 1e-13 * np.abs(expect).max()


def test_interface_mass_is_permuted_face_mass(geo):
    m, g = geo
    V = m.volumes()
    checked = 0
    for e in range(m.NE):
        for f in range(4):
            ep = m.nbr[e, f]
            if ep < 0:
                continue
            c = 5.0 * g["area"][e, f] / (3.0 * V[e])
            E = np.zeros((4, 4))
            for b in FN[f]:
                bp = int(np.nonzero(m.node[ep] == m.node[e, b])[0][0])
                for a in range(4):
                    E[a, bp] = int(a == b) + int(a in FN[f]) - 0.8
            assert np.abs(g["IIM"][e, f] - c * E).max() <= 1e-13 * c
            checked += 1
    assert checked == int((m.nbr >= 0).sum())


def test_normals_unit_outward_and_closed(geo):
    m, g = geo
    n, A = g["nrm"], g["area"]
    assert np.allclose(np.linalg.norm(n, axis=2), 1.0, atol=1e-15)
    cen = m.centroids()
    for f in range(4):
        fc = m.rg[m.node[:, FN[f]]].mean(axis=1)
        assert np.all(np.einsum("ij,ij->i", n[:, f], fc - cen) > 0)
    closure = np.einsum("ef,efc->ec", A, n)
    assert np.abs(closure).max() <= 1e-14 * A.max()


def test_normals_neighbour_exact_negation(geo):
    m, g = geo
    n = g["nrm"]
    for e in range(m.NE):
        for f in range(4):
            ep = m.nbr[e, f]
            if ep < 0:
                continue
            fp = int(np.nonzero(m.nbr[ep] == e)[0][0])
            assert np.array_equal(n[e, f], -n[ep, fp])


def test_gradient_normal_identity(geo):
    """grad(lambda_a) = -(A_opp / 3V) n_opp: the compact form the HIP sweep uses."""
    m, g = geo
    grad = _grads(m)
    V = m.volumes()
    for a in range(4):
        f = int(np.nonzero(OPP == a)[0][0])
        expect = -(g["area"][:, f] / (3 * V))[:, None] * g["nrm"][:, f]
        assert np.abs(grad[:, a] - expect).max() <= 1e-12 * np.abs(grad).max()


def test_neighbors_match_sorted_face_convention(small_mesh):
    m = small_mesh
    for e in range(m.NE):
        for f in range(4):
            ep = m.nbr[e, f]
            mine = sorted(m.node[e, FN[f]].tolist())
            if ep < 0:
                continue
            theirs = [sorted(m.node[ep, FN[g]].tolist()) for g in range(4)]
            assert mine in theirs
    # boundary face count of an n^3 Kuhn cube: 6 faces * n^2 squares * 2 triangles
    assert int((m.nbr < 0).sum()) == 6 * 3 * 3 * 2
