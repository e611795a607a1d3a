"""The persistent dataflow sweep (one kernel, workgroups taking work in level
order and waiting on per-(direction, cell) completion flags) produces the same
flux as the level-by-level launches, bit for bit, and reports a failed wait
instead of hanging."""
import numpy as np
import pytest

from lionbolt_amd import _lib, angular, mesh as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("G", [1, 5, 24])
def test_persistent_equals_level_launches(G):
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(6, a=5.0, seed=1007, slab=(2.0, 3.0))
    ang = angular.product(4, 8, 1)
    op = Operator(m, ang)
    sigt = np.stack([np.linspace(0.1, 1.5, G), np.linspace(0.3, 4.5, G)], axis=1)
    op.set_physics_groups(sigt, np.zeros((G, 2, ang.L + 1)))
    q = torch.rand(ang.NI * m.NE * 4 * G, dtype=torch.float64, device="cuda")
    res = {}
    for mode in ("levels", "persistent", "persistent"):
        op.set_sweep_mode(mode)
        psi = torch.full_like(q, float("nan"))
        op.transport_matinv_dev(q, psi, 0, G)
        op.synchronize()
        res.setdefault(mode, []).append(psi.cpu().numpy())
    for v in res["persistent"]:
        assert np.array_equal(v, res["levels"][0])


def test_persistent_solver_matches_levels():
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(5, a=4.0)
    ang = angular.snprep(3)
    op = Operator(m, ang)
    op.set_physics([1.0], [[0.8, 0.3, 0.1, 0.05]])
    q = np.random.default_rng(3).random(op.n)
    op.set_sweep_mode("levels")
    a, sa = op.boltzmann_matinv(q, space="moment", m=30)
    op.set_sweep_mode("persistent")
    b, sb = op.boltzmann_matinv(q, space="moment", m=30)
    assert sa["iterates"] == sb["iterates"]
    assert np.array_equal(a, b)


def test_persistent_wait_limit_reports_failure():
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(8, a=4.0)
    ang = angular.snprep(3)
    op = Operator(m, ang)
    op.set_physics([1.0], [[0.0] * 4])
    op.set_sweep_mode("persistent")
    op.set_sweep_wait_limit(0)            # every lane that finds its upstream unfinished gives up at once
    q = torch.rand(op.n, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(q)
    op.transport_matinv_dev(q, psi, 0, 1)
    with pytest.raises(_lib.LionboltError) as ei:
        op.synchronize()
    assert "wait limit" in str(ei.value)
    op.set_sweep_wait_limit(1 << 22)
    op.transport_matinv_dev(q, psi, 0, 1)
    op.synchronize()
    op.set_sweep_mode("levels")
    ref = torch.empty_like(q)
    op.transport_matinv_dev(q, ref, 0, 1)
    op.synchronize()
    assert torch.equal(psi, ref)
