"""The sweep-order flux layout of the within-group solves gives the same numbers
as the natural layout, bit for bit.

* A one-group sweep that writes each direction's flux at its cells' sweep
  positions, mapped back to the natural layout, equals the natural sweep in
  every launch mode (level launches, graph replay, direction-resident and LDS
  kernels at each workgroup size), from a natural source and, in place, from a
  source in sweep order.  The LDS plan's staged and global trace reads, which
  address the neighbour's flux by its sweep position, run at the full 384k-tet
  size (tests/test_gpu_full_size.py).
* M S phi written in sweep order holds the natural product's values, and D psi
  gathered from the sweep-order flux equals D psi of the natural flux.
* Multigroup flux-moment GMRES and FGMRES, and the single-group flux-moment
  MatInv, give the same moments, iterates and angular flux in all three
  layouts.
* On a problem set the persistent sweep's sweep-order flux (the batch's
  problem-major order) and the batched GMRES agree with the natural layout
  (the true residuals, evaluated by two different exact formulas, to 1e-3)."""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M, problems

pytestmark = pytest.mark.gpu

MODES = [("levels", 0), ("graph", 0), ("direction", 256), ("direction", 512), ("direction", 768),
         ("direction", 1024), ("direction_lds", 256), ("direction_lds", 512), ("direction_lds", 768)]


def _sweep_both(op, q, modes):
    """The natural sweep of q against its sweep-order sweeps in each launch mode, mapped
    back to the natural layout; returns the natural flux and the last sweep-order one."""
    import torch
    ref = torch.full_like(q, float("nan"))
    op.set_sweep_mode("levels")
    op.transport_matinv_dev(q, ref, 0, 1)
    op.synchronize()
    for mode, block in modes:
        op.set_sweep_mode(mode)
        if block:
            op.set_direction_block(block)
        so = torch.full_like(q, float("nan"))
        op.transport_matinv_order_dev(q, so, 0)
        op.synchronize()
        nat = torch.full_like(q, float("nan"))
        op.flux_to_natural_dev(so, nat)
        assert torch.equal(nat, ref), (mode, block)
        assert not torch.equal(so, ref), (mode, block)   # the layouts do differ
    return ref, so


def _to_order(op, v):
    """v (natural layout) in sweep order, through the positions the library uses:
    flux_to_natural_dev is a permutation, so its inverse is read off an index vector."""
    import torch
    idx = torch.arange(v.numel(), dtype=torch.float64, device=v.device)
    perm = torch.empty_like(idx)
    op.flux_to_natural_dev(idx, perm)          # perm[natural slot] = sweep-order slot
    out = torch.empty_like(v)
    out[perm.long()] = v
    return out


def test_sweep_order_flux_equals_natural_sweep():
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(6, a=5.0, seed=1007, slab=(2.0, 3.0))
    ang = angular.product(4, 8, 3)
    op = Operator(m, ang)
    op.set_physics_groups(np.array([[0.4, 1.7]]), np.zeros((1, 2, ang.L + 1)))
    q = torch.rand(ang.NI * m.NE * 4, dtype=torch.float64, device="cuda")
    ref, so = _sweep_both(op, q, MODES)
    phi_nat = torch.empty(op.nmom * op.NENK, dtype=torch.float64, device="cuda")
    phi_so = torch.empty_like(phi_nat)
    op.moments_dev(ref, phi_nat)
    op.moments_order_dev(so, phi_so)
    assert torch.equal(phi_so, phi_nat)
    with pytest.raises(Exception):
        op.transport_matinv_order_dev(q, q, 0)   # a natural source cannot share the flux's buffer
    # the source in sweep order too, swept in place
    q_so = _to_order(op, q)
    for mode, block in MODES:
        op.set_sweep_mode(mode)
        if block:
            op.set_direction_block(block)
        v = q_so.clone()
        op.transport_matinv_order_dev(v, v, 0, source_in_order=True)
        op.synchronize()
        assert torch.equal(v, so), (mode, block)
    # M S phi in sweep order holds the natural product's values
    phi = torch.rand_like(phi_nat)
    m_nat = torch.empty_like(q)
    m_so = torch.full_like(q, float("nan"))
    op.moments_to_discrete_dev(phi, m_nat, 0, 1)
    op.moments_to_discrete_order_dev(phi, m_so, 0)
    assert torch.equal(m_so, _to_order(op, m_nat))
    op.close()


@pytest.mark.parametrize("kind", ["gmres", "fgmres"])
def test_multigroup_solve_same_in_both_layouts(kind):
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(5, a=4.0, seed=1021, slab=(1.5, 2.5))
    ang = angular.product(4, 8, 3)
    G = 3
    sigt, S = problems.downscatter_xs(G=G, L=ang.L)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, problems.within_group(S))
    if kind == "fgmres":
        from lionbolt_amd import learned_pc
        params = learned_pc.init_params(seed=3)
        op.set_preconditioner(params)
    nphi = op.nmom * op.NENK
    g = torch.Generator(device="cuda").manual_seed(11)
    qmom = torch.rand(G * nphi, dtype=torch.float64, device="cuda", generator=g)
    out = {}
    for order in ("natural", "sweep", "sweep_source"):
        op.set_flux_order(order)
        phi = torch.empty_like(qmom)
        stats = op.multigroup_solve_dev(S, phi, qmom=qmom, kind=kind, space="moment", m=10, rtol=1e-10)
        out[order] = (phi.cpu().numpy(), [s["iterates"] for s in stats], [s["true_residual"] for s in stats])
    for order in ("sweep", "sweep_source"):
        assert np.array_equal(out[order][0], out["natural"][0]), order
        assert out[order][1] == out["natural"][1], order
        np.testing.assert_allclose(out[order][2], out["natural"][2], rtol=1e-10, atol=1e-300)
    op.close()


def test_single_group_moment_matinv_same_in_both_layouts():
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(5, a=4.0, seed=1031)
    ang = angular.product(4, 8, 2)
    op = Operator(m, ang)
    op.set_physics([0.8], [[0.7, 0.2, 0.05]])
    q = np.random.default_rng(5).random(op.n)
    out = {}
    for order in ("natural", "sweep", "sweep_source"):
        op.set_flux_order(order)
        x, st = op.boltzmann_matinv(q, kind="gmres", space="moment", m=20, rtol=1e-10)
        out[order] = (x, st["iterates"])
    for order in ("sweep", "sweep_source"):
        assert np.array_equal(out[order][0], out["natural"][0]), order
        assert out[order][1] == out["natural"][1], order
    op.close()


def test_problem_set_sweep_order_flux():
    """A problem set (five uneven cubes, S8): the persistent sweep writing the flux in
    the batch's problem-major sweep order equals the natural persistent sweep mapped
    back, at each persistent configuration; the batched GMRES gives the same moments
    and iterates with the flux in sweep order as in the natural layout."""
    import torch
    from test_gpu_batch import batch_op, small_batch
    um, eoff, sigt, sigs, probs = small_batch()
    ang = angular.level_symmetric_s8()
    op = batch_op(um, eoff, sigt, sigs, ang)
    q = torch.rand(op.n, dtype=torch.float64, device="cuda")
    op.set_sweep_mode("persistent")
    for block, per_cu in ((256, 4), (512, 2), (1024, 2)):
        op.set_direction_block(block)
        op.set_workgroups_per_cu(per_cu)
        ref = torch.full_like(q, float("nan"))
        op.transport_matinv_dev(q, ref, 0, 1)
        for _ in range(2):                       # graph replays with eager work between
            so = torch.full_like(q, float("nan"))
            op.transport_matinv_order_dev(q, so, 0)
            op.synchronize()
            nat = torch.full_like(q, float("nan"))
            op.flux_to_natural_dev(so, nat)
            assert torch.equal(nat, ref), (block, per_cu)
    op.set_sweep_mode("levels")
    with pytest.raises(Exception):               # positions follow the persistent kernel's order
        op.transport_matinv_order_dev(q, so, 0)
    op.set_sweep_mode("auto")
    qmom = problems.uneven_beams_sources(op, eoff, probs, sigt, sigs).reshape(-1).contiguous()
    out = {}
    for order in ("natural", "sweep"):
        op.set_flux_order(order)
        phi = torch.empty(op.nmom * op.NENK, dtype=torch.float64, device="cuda")
        stats = op.batch_solve_dev(qmom, phi, m=30, rtol=1e-10)
        out[order] = (phi.cpu().numpy(), [s["iterates"] for s in stats], [s["true_residual"] for s in stats])
    assert np.array_equal(out["sweep"][0], out["natural"][0])
    assert out["sweep"][1] == out["natural"][1]
    # the sweep-order check evaluates T^-1 M S (D psi - phi), the natural one T^-1 q - (I - T^-1 K) psi
    np.testing.assert_allclose(out["sweep"][2], out["natural"][2], rtol=1e-3)
    assert max(out["sweep"][2]) <= 1e-9
    assert op.sweep_order_choice()[0] == "persistent"
    op.close()


def test_fp32_flux_output_is_the_rounded_fp64_sweep():
    """The FP32 flux mode's sweep: its FP32 output is the FP64 sweep-order flux
    rounded to nearest, bit for bit, at each LDS workgroup size (the arithmetic
    and the LDS traces stay FP64; the FP64 flux goes to global memory only for
    cells read from there, and any such read of a value not written would change
    the fluxes downstream).  D psi of the FP32 flux equals the FP64 gather on the
    same values widened to FP64, bit for bit."""
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(6, a=5.0, seed=1007, slab=(2.0, 3.0))
    ang = angular.product(4, 8, 3)
    op = Operator(m, ang)
    op.set_physics_groups(np.array([[0.4, 1.7]]), np.zeros((1, 2, ang.L + 1)))
    q = torch.rand(ang.NI * m.NE * 4, dtype=torch.float64, device="cuda")
    for block in (256, 512, 768):
        op.set_direction_block(block)
        op.set_sweep_mode("direction_lds")
        ref = torch.empty_like(q)
        op.transport_matinv_order_dev(q, ref, 0)
        so = torch.full_like(q, float("nan"))
        f32 = torch.full((q.numel(),), float("nan"), dtype=torch.float32, device="cuda")
        op.transport_matinv_order32_dev(q, so, f32, 0)
        op.synchronize()
        assert torch.equal(f32, ref.float()), block
    phi32 = torch.empty(op.nmom * op.NENK, dtype=torch.float64, device="cuda")
    phi64 = torch.empty_like(phi32)
    op.moments_order32_dev(f32, phi32)
    op.moments_order_dev(f32.double(), phi64)
    assert torch.equal(phi32, phi64)
    op.close()


def test_fp32_flux_mode_multigroup_solve():
    """The multigroup flux-moment GMRES with the Arnoldi steps' flux in FP32: the
    true residual (FP64 sweeps) meets the tolerance and the moments agree with the
    FP64 solve to well within it."""
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(5, a=4.0, seed=1021, slab=(1.5, 2.5))
    ang = angular.product(4, 8, 3)
    G = 3
    sigt, S = problems.downscatter_xs(G=G, L=ang.L)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, problems.within_group(S))
    nphi = op.nmom * op.NENK
    g = torch.Generator(device="cuda").manual_seed(11)
    qmom = torch.rand(G * nphi, dtype=torch.float64, device="cuda", generator=g)
    out = {}
    for order in ("sweep", "sweep_fp32"):
        op.set_flux_order(order)
        phi = torch.empty_like(qmom)
        stats = op.multigroup_solve_dev(S, phi, qmom=qmom, kind="gmres", space="moment", m=30, rtol=1e-6)
        out[order] = (phi.view(G, -1).cpu().numpy(), stats)
    for g_ in range(G):
        a, b = out["sweep_fp32"][0][g_], out["sweep"][0][g_]
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max(), g_
        assert out["sweep_fp32"][1][g_]["true_residual"] <= 1e-6, g_
        assert abs(out["sweep_fp32"][1][g_]["iterates"] - out["sweep"][1][g_]["iterates"]) <= 1, g_
    op.close()
