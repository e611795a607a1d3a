"""The allocation-free network apply (lb_mlp_apply_noalloc_dev) against the
allocating one (lb_mlp_apply_dev), through the C ABI: two calls in a row, the
second on a smaller operator, in both network formats.  The two entry points
share their kernels and compute the residual's scale with the same reduction
and the same IEEE arithmetic, so the outputs are expected to agree bit for bit;
the bound written here is the FP32 path's own parity bar (1e-5 relative)."""
import numpy as np
import pytest

from lionbolt_amd import angular, learned_pc, mesh as M

pytestmark = pytest.mark.gpu


def _op(n, seed):
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(n, a=6.0, seed=seed, slab=(2.0, 3.5))
    ang = angular.product(2, 8, 3)
    op = Operator(m, ang)
    sigt = np.array([0.4, 1.2])
    sigs = np.stack([0.99 * sigt[0] * 0.6 ** np.arange(4), 0.99 * sigt[1] * 0.6 ** np.arange(4)])
    op.set_physics(sigt, sigs)
    return op, m


@pytest.mark.parametrize("fmt", ["bf16", "fp32"])
def test_noalloc_apply_equals_allocating_apply(fmt):
    import torch
    params = learned_pc.init_params(11)
    for n, seed in ((6, 301), (3, 302)):            # the second call has the smaller input
        op, m = _op(n, seed)
        op.set_preconditioner(params, fmt=fmt)
        r = torch.randn(16 * m.NENK, dtype=torch.float64, device="cuda")
        ref = torch.full_like(r, float("nan"))
        op.mlp_apply_dev(r, ref)
        got = torch.full_like(r, float("nan"))
        op.mlp_apply_noalloc_dev(r, got)
        assert bool(torch.isfinite(got).all())
        diff = float((got - ref).abs().max() / ref.abs().max())
        print(f"noalloc {fmt} n={n}: max rel diff {diff:.3e}", flush=True)
        assert diff <= 1e-5
        op.close()


def test_noalloc_apply_refuses_without_a_network():
    import torch
    from lionbolt_amd import _lib
    op, m = _op(2, 303)
    r = torch.randn(16 * m.NENK, dtype=torch.float64, device="cuda")
    z = torch.empty_like(r)
    with pytest.raises(_lib.LionboltError) as ei:
        op.mlp_apply_noalloc_dev(r, z)
    assert ei.value.code == -3
    op.close()
