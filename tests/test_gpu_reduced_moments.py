"""The flux-moment GMRES with its Arnoldi steps' D psi product on MFMA in a
reduced input format (FP32 accumulation), the right-hand side, psi and the
true residual in FP64.  The reduced product perturbs K by about its input
rounding (split BF16: ~1e-5 of the moments), so the solve lands within that
distance of the FP64 solution and its FP64 true residual says how far."""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M, problems

pytestmark = pytest.mark.gpu


def _setup(G=3):
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(5, a=10.0, seed=1001, slab=(4.0, 6.0))
    ang = problems.product_8x32()                       # 256 directions, P3
    sigt, S = problems.downscatter_xs(G=G, L=ang.L)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, problems.within_group(S))
    qmom = problems.beam_group_sources(op, m, S, sigt)[0]
    return op, S, qmom, torch


def test_split_bf16_d_psi_solve_tracks_fp64():
    op, S, qmom, torch = _setup()
    phi = torch.zeros_like(qmom)
    st = op.multigroup_solve_dev(S, phi, qmom=qmom, kind="gmres", space="moment", m=30, rtol=1e-6)
    assert all(s["status"] == 0 for s in st)
    assert max(s["true_residual"] for s in st) <= 1e-6
    op.set_moment_precision("split_bf16")
    phr = torch.zeros_like(qmom)
    sr = op.multigroup_solve_dev(S, phr, qmom=qmom, kind="gmres", space="moment", m=30, rtol=1e-6, maxit=60,
                                 raise_on_fail=False)
    res = [s["true_residual"] for s in sr]
    assert all(np.isfinite(res)) and max(res) <= 1e-3, res
    ref = phi[:, 0]
    diff = float(((phr[:, 0] - ref).abs().amax(dim=1) / ref.abs().amax(dim=1)).max())
    assert 0.0 < diff <= 1e-3, diff                     # perturbed, and only by about the rounding
    op.set_moment_precision("fp64")                      # back to FP64: the same iterates as the first solve
    ph2 = torch.zeros_like(qmom)
    s2 = op.multigroup_solve_dev(S, ph2, qmom=qmom, kind="gmres", space="moment", m=30, rtol=1e-6)
    assert [s["iterates"] for s in s2] == [s["iterates"] for s in st]
    assert torch.equal(ph2, phi)
    op.close()


def test_reduced_precision_needs_256_directions_and_p3():
    from lionbolt_amd import _lib
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(3, a=4.0, seed=1001)
    op = Operator(m, angular.product(2, 8, 1))
    with pytest.raises(_lib.LionboltError):
        op.set_moment_precision("split_bf16")
    op.set_moment_precision("fp64")
    op.close()
