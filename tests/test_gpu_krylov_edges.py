"""Edge cases of the flux-moment GMRES (Arnoldi on the scattering-sweep
operator K with fused classical Gram-Schmidt), against the oracle's GMRES on
the same problem and against closed forms:

* restarts: GMRES(3) reaches the GMRES(30) solution (both to rtol 1e-10;
  scalar flux within 1e-7 relative of the oracle);
* the iteration limit returns LB_E_MAXIT with a finite iterate;
* no scattering (K = 0): the Krylov space stops growing after one step, and
  the solution is the sweep of the source, psi = T^-1 q, bit for bit, with a
  true residual of exactly zero;
* a zero source gives a zero flux without an Arnoldi step."""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M
from oracle.oracle import OracleOp

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def problem():
    m = M.jittered_cube(5, a=4.0, seed=1017, slab=(1.5, 2.5))
    ang = angular.snprep(3)
    q = np.random.default_rng(5).random(4 * m.NE * ang.NI)
    return m, ang, q


def test_restarted_moment_gmres_matches_oracle(problem):
    from lionbolt_amd.ops import Operator
    m, ang, q = problem
    sigs = np.array([[0.7, 0.3, 0.1, 0.05], [2.1, 0.9, 0.3, 0.15]])
    op = Operator(m, ang)
    op.set_physics([1.0, 3.0], sigs)
    a, sa = op.boltzmann_matinv(q, space="moment", m=30, rtol=1e-10)
    b, sb = op.boltzmann_matinv(q, space="moment", m=3, rtol=1e-10)
    assert sa["status"] == 0 and sb["status"] == 0
    assert sb["matvecs"] > sa["matvecs"]            # restarts cost extra operator applications
    assert sa["true_residual"] <= 1e-9 and sb["true_residual"] <= 1e-9
    orc = OracleOp(m, ang)
    orc.set_physics([1.0, 3.0], sigs)
    c, sc = orc.boltzmann_matinv(q, kind="gmres", m=30, rtol=1e-10)
    assert sc["status"] == 0
    ref = orc.fluence(c)
    assert rel(op.fluence(a), ref) <= 1e-7
    assert rel(op.fluence(b), ref) <= 1e-7


def test_iteration_limit_returns_maxit(problem):
    from lionbolt_amd import _lib
    from lionbolt_amd.ops import Operator
    m, ang, q = problem
    op = Operator(m, ang)
    op.set_physics([1.0, 3.0], np.array([[0.95, 0.3, 0.1, 0.05], [2.9, 0.9, 0.3, 0.15]]))
    psi, st = op.boltzmann_matinv(q, space="moment", m=30, rtol=1e-14, maxit=2, raise_on_fail=False)
    assert _lib.ERRORS[st["status"]] == "LB_E_MAXIT"
    assert st["matvecs"] <= 2 and np.all(np.isfinite(psi))


def test_no_scattering_gives_the_sweep_of_the_source(problem):
    from lionbolt_amd.ops import Operator
    m, ang, q = problem
    op = Operator(m, ang)
    op.set_physics([1.0, 3.0], np.zeros((2, 4)))
    psi, st = op.boltzmann_matinv(q, space="moment", m=30, rtol=1e-10)
    assert st["status"] == 0 and st["true_residual"] == 0.0
    t = op.transport_matinv(q)
    assert np.array_equal(psi, t)


def test_zero_source_gives_zero_flux(problem):
    from lionbolt_amd.ops import Operator
    m, ang, _ = problem
    op = Operator(m, ang)
    op.set_physics([1.0, 3.0], np.array([[0.7, 0.3, 0.1, 0.05], [2.1, 0.9, 0.3, 0.15]]))
    psi, st = op.boltzmann_matinv(np.zeros(4 * m.NE * ang.NI), space="moment", m=30, rtol=1e-10)
    assert st["status"] == 0 and st["matvecs"] == 0
    assert not np.any(psi)
