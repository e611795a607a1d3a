"""Pins for the oracle's transport inverse T^-1 (TransportAlgebra.f90:29-246).

* dense solve: the DFEM system of one direction assembled as a full matrix and
  solved directly equals the ordered sweep (checks ordering, upstream addressing
  and the encode/decode permutations);
* particle balance: for every direction, sigma * integral(psi) + boundary
  outflow = integral(q), a closed-form consequence of the weak form that holds
  to rounding for the upwind discretisation;
* TransportMMS (validation/operators/TransportMMS.f90:44-89): psi = exp(-A r^2)
  with source (-2A k.r + sigma) exp(-A r^2) converges at second order under
  refinement of the jittered cube on [-1, 1]^3.
"""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M
from oracle.oracle import OracleOp

FN = M.FACE_NODES
EPS = np.finfo(float).eps


def test_sweep_equals_dense_solve(small_mesh):
    m = small_mesh
    ang = angular.snprep(1)
    op = OracleOp(m, ang)
    sigt = 0.7
    op.set_physics([sigt], [[0.0, 0.0]])
    g = op.geometry()
    rng = np.random.default_rng(5)
    q = rng.standard_normal(op.n)
    psi = op.transport_matinv(q).reshape(ang.NI, m.NE, 4)
    for i in range(ang.NI):
        k = ang.k[i]
        A = np.zeros((m.NENK, m.NENK))
        kn = np.einsum("efc,c->ef", g["nrm"], k)
        for e in range(m.NE):
            T = sigt * np.eye(4) - np.einsum("d,dab->ab", k, g["IC"][e])
            for f in range(4):
                if kn[e, f] >= 0:
                    T += kn[e, f] * g["IFM"][e, f]
                ep = m.nbr[e, f]
                if ep >= 0 and kn[e, f] < -EPS:
                    A[4 * e:4 * e + 4, 4 * ep:4 * ep + 4] = kn[e, f] * g["IIM"][e, f]
            A[4 * e:4 * e + 4, 4 * e:4 * e + 4] = T
        x = np.linalg.solve(A, q.reshape(ang.NI, -1)[i])
        assert np.abs(x - psi[i].reshape(-1)).max() <= 1e-12 * np.abs(x).max()


def test_particle_balance(small_mesh):
    m = small_mesh
    ang = angular.level_symmetric_s8()
    op = OracleOp(m, ang)
    sigt = 0.3
    op.set_physics([sigt], [[0.0]])
    g = op.geometry()
    V = m.volumes()
    rng = np.random.default_rng(7)
    q = rng.random(op.n)
    psi = op.transport_matinv(q).reshape(ang.NI, m.NE, 4)
    qq = q.reshape(ang.NI, m.NE, 4)
    for i in range(ang.NI):
        kn = np.einsum("efc,c->ef", g["nrm"], ang.k[i])
        absorb = sigt * np.sum(V / 4 * psi[i].sum(axis=1))
        src = np.sum(V / 4 * qq[i].sum(axis=1))
        leak = 0.0
        for f in range(4):
            bo = (m.nbr[:, f] < 0) & (kn[:, f] >= 0)
            leak += np.sum(kn[bo, f] * g["area"][bo, f] / 3 * psi[i][bo][:, FN[f]].sum(axis=1))
        assert abs(absorb + leak - src) <= 1e-12 * src


def test_zero_source_and_linearity(small_mesh):
    m = small_mesh
    ang = angular.snprep(1)
    op = OracleOp(m, ang)
    op.set_physics([1.3], [[0.0, 0.0]])
    assert np.all(op.transport_matinv(np.zeros(op.n)) == 0)
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(op.n), rng.standard_normal(op.n)
    lhs = op.transport_matinv(2.0 * a - 3.0 * b)
    rhs = 2.0 * op.transport_matinv(a) - 3.0 * op.transport_matinv(b)
    assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(rhs).max()


def mms_error(n, L=3):
    m = M.jittered_cube(n, a=2.0, seed=1001)
    m.rg -= 1.0
    ang = angular.snprep(L)
    op = OracleOp(m, ang)
    A, sig = 10.0, 1e-10
    op.set_physics([sig], [[0.0] * (L + 1)])
    r = m.dof_coords()
    r2 = np.einsum("jc,jc->j", r, r)
    kr = ang.k @ r.T                              # (NI, NENK)
    s = (-2.0 * A * kr + sig) * np.exp(-A * r2)[None, :]
    exact = np.broadcast_to(np.exp(-A * r2), s.shape)
    psi = op.transport_matinv(s).reshape(s.shape)
    return np.sqrt(np.mean((psi - exact) ** 2))


def test_transport_mms_second_order():
    e1, e2 = mms_error(6), mms_error(12)
    assert e2 < e1
    assert e1 / e2 > 3.0     # second order: ~4x per halving of h
