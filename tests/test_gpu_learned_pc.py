"""Learned right preconditioner: MLP inference (BF16 MFMA and FP32) against an
FP64 numpy forward pass, training gradients against torch autograd, the
solver-generated training data, and flexible GMRES with and without the
stagnation fallback (checked against plain flux-moment GMRES)."""
import math

import numpy as np
import pytest

from lionbolt_amd import _lib, angular, learned_pc, mesh as M, problems

pytestmark = pytest.mark.gpu


def small_op(n=4, c=0.99, st=0.4):
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(n, a=6.0, seed=77, slab=(2.0, 3.5))
    ang = angular.product(4, 8, 3)
    op = Operator(m, ang)
    sigt = np.array([st, 3 * st])
    sigs = np.stack([c * sigt[0] * 0.6 ** np.arange(4), c * sigt[1] * 0.6 ** np.arange(4)])
    op.set_physics(sigt, sigs)
    return op, m, ang, sigt, sigs


def test_mlp_inference_matches_fp64_forward():
    import torch
    op, m, ang, sigt, sigs = small_op()
    params = learned_pc.init_params(7)
    params[learned_pc.SHAPES[0][1][0] * 64:][:64] = 0.1            # nonzero b1
    r = torch.randn(16 * m.NENK, dtype=torch.float64, device="cuda")
    feat, s = op.mlp_features_dev(r)
    X = feat.cpu().numpy().astype(np.float64)
    # features: normalised moments, log(sigma_t h), c
    rn = r.cpu().numpy().reshape(16, -1).T
    assert np.allclose(X[:, :16], rn / s, rtol=1e-6, atol=1e-6)
    assert abs(s - math.sqrt((rn ** 2).mean())) <= 1e-12 * s
    h = np.cbrt(6.0 * m.volumes())
    st_e = sigt[m.mat - 1]
    assert np.allclose(X[:, 16], np.repeat(np.log(st_e * h), 4), rtol=1e-6, atol=1e-6)
    assert np.allclose(X[:, 17], 0.99, rtol=1e-6)
    ref = learned_pc.forward_reference(params.astype(np.float64), X) * s
    for fmt, tol in (("fp32", 1e-5), ("bf16", 3e-2)):
        op.set_preconditioner(params, fmt)
        z = torch.empty_like(r)
        op.mlp_apply_dev(r, z)
        got = z.cpu().numpy().reshape(16, -1).T
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err <= tol, (fmt, err)


def test_training_gradient_matches_autograd():
    import torch
    g = torch.Generator().manual_seed(3)
    B = 384
    X = torch.randn(B, 18, generator=g)
    Y = torch.randn(B, 16, generator=g)
    params = learned_pc.init_params(11)
    grad = np.empty_like(params)
    loss = np.empty(1, dtype=np.float32)
    Xd, Yd = X.cuda(), Y.cuda()                 # keep the device copies alive across the call
    _lib.check(_lib.lib().lb_mlp_grad_dev(Xd.data_ptr(), Yd.data_ptr(), B,
                                          params.ctypes.data_as(_lib._f), params.size,
                                          grad.ctypes.data_as(_lib._f), loss.ctypes.data_as(_lib._f)))
    P = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in learned_pc.unpack(params).items()}
    h = torch.tanh(X.double() @ P["W1"] + P["b1"])
    h = torch.tanh(h @ P["W2"] + P["b2"])
    h = torch.tanh(h @ P["W3"] + P["b3"])
    out = h @ P["W4"] + P["b4"]
    L = ((out - Y.double()) ** 2).mean()
    L.backward()
    ref = np.concatenate([P[k].grad.numpy().reshape(-1) for k, _ in learned_pc.SHAPES])
    assert abs(loss[0] - L.item()) <= 1e-5 * L.item()
    assert np.abs(grad - ref).max() <= 1e-4 * np.abs(ref).max()


def test_training_fits_a_smooth_map():
    import torch
    g = torch.Generator().manual_seed(5)
    N = 20000
    X = torch.randn(N, 18, generator=g)
    A = torch.randn(18, 16, generator=g) / 4
    Y = torch.tanh(X @ A) + 0.1 * X[:, :16]
    params, info = learned_pc.train(X.cuda(), Y.cuda(), steps=1500, batch=1024, lr=2e-3)
    assert info["loss_last"] < 0.1 * info["loss_first"]
    pred = learned_pc.forward_reference(params.astype(np.float64), X.numpy().astype(np.float64))
    assert ((pred - Y.numpy()) ** 2).mean() < 0.1 * (Y.numpy() ** 2).mean()


def test_snapshots_residuals_and_final_iterate():
    import torch
    op, m, ang, sigt, sigs = small_op(c=0.9)
    qm = torch.zeros(16 * m.NENK, dtype=torch.float64, device="cuda")
    qm[:m.NENK] = math.sqrt(4 * math.pi)
    X, Y, st, cnt = learned_pc.collect(op, qm, m=30, rtol=1e-10, rows_per_snap=64)
    assert st["status"] == 0 and cnt == st["iterates"] and len(X) >= cnt - 1
    # GMRES residuals do not grow inside a cycle: the target scale falls
    ymax = [float(y.abs().max()) for y in Y]
    assert all(np.isfinite(ymax))


@pytest.mark.parametrize("fmt", ["bf16", "fp32"])
def test_fgmres_with_learned_pc_reaches_the_gmres_solution(fmt):
    import torch
    op, m, ang, sigt, sigs = small_op(n=5)
    qm = torch.zeros(16 * m.NENK, dtype=torch.float64, device="cuda")
    qm[:m.NENK] = math.sqrt(4 * math.pi)
    X, Y, _, _ = learned_pc.collect(op, qm, rows_per_snap=256)
    params, info = learned_pc.train(torch.cat(X), torch.cat(Y), steps=600, batch=1024)
    op.set_preconditioner(params, fmt)
    q = torch.empty(op.n, dtype=torch.float64, device="cuda")
    op.discrete_source_dev(qm, q)
    qh = q.cpu().numpy()
    a, sa = op.boltzmann_matinv(qh, kind="gmres", space="moment", m=30, rtol=1e-9)
    b, sb = op.boltzmann_matinv(qh, kind="fgmres", space="moment", m=30, rtol=1e-9)
    assert sa["status"] == 0 and sb["status"] == 0
    assert sb["true_residual"] <= 1e-7
    pa, pb = op.fluence(a), op.fluence(b)
    assert np.abs(pa - pb).max() <= 1e-6 * np.abs(pa).max()


def test_fgmres_falls_back_when_the_network_returns_nothing():
    import torch
    op, m, ang, sigt, sigs = small_op(n=4, c=0.95)
    params = np.zeros(_lib.MLP_PARAMS, dtype=np.float32)     # P(r) = 0: unusable
    op.set_preconditioner(params, "bf16")
    q = np.random.default_rng(2).random(op.n)
    a, sa = op.boltzmann_matinv(q, kind="gmres", space="moment", m=30, rtol=1e-9)
    b, sb = op.boltzmann_matinv(q, kind="fgmres", space="moment", m=30, rtol=1e-9)
    assert sb["status"] == 0 and sb["fallback"] == 1
    assert np.abs(op.fluence(a) - op.fluence(b)).max() <= 1e-7 * np.abs(op.fluence(a)).max()


def test_preconditioner_argument_checks():
    op, m, ang, sigt, sigs = small_op(n=2)
    with pytest.raises(_lib.LionboltError):
        op.set_preconditioner(np.zeros(10, dtype=np.float32))
    bad = np.zeros(_lib.MLP_PARAMS, dtype=np.float32)
    bad[5] = np.nan
    with pytest.raises(_lib.LionboltError):
        op.set_preconditioner(bad)
    q = np.ones(op.n)
    with pytest.raises(_lib.LionboltError) as ei:
        from lionbolt_amd.ops import Operator
        o2 = Operator(m, ang)
        o2.set_physics(sigt, sigs)
        o2.boltzmann_matinv(q, kind="fgmres", space="moment")
    assert ei.value.code == -3


def test_threshold_chooses_plain_gmres_per_group():
    """lb_op_set_preconditioner_threshold: groups thinner than the threshold run
    the plain flux-moment GMRES (fallback -1, the same flux bit for bit); an
    infinite threshold turns the network off, zero leaves it on everywhere."""
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(4, a=6.0, seed=77, slab=(2.0, 3.5))
    ang = angular.product(4, 8, 3)
    G = 3
    sigt = np.array([[0.05, 0.15], [0.4, 1.2], [1.5, 4.5]])
    sigs = np.stack([np.stack([0.99 * s * 0.6 ** np.arange(4) for s in row]) for row in sigt])
    S = np.zeros((G, G, 2, 4))
    for g in range(G):
        S[g, g] = sigs[g]
    op = Operator(m, ang)
    op.set_physics_groups(sigt, sigs)
    op.set_preconditioner(learned_pc.init_params(5), "bf16")
    qm = torch.zeros((G, 16, m.NENK), dtype=torch.float64, device="cuda")
    qm[:, 0] = math.sqrt(4 * math.pi)
    plain = torch.zeros_like(qm)
    op.multigroup_solve_dev(S, plain, qmom=qm, kind="gmres", rtol=1e-9)
    thick = np.array([sigt[g].max() * np.cbrt(m.volumes().sum()) for g in range(G)])
    mid = float(0.5 * (thick[0] + thick[1]))              # between the first two groups' thicknesses
    for thr, used in ((math.inf, [False] * G), (0.0, [True] * G), (mid, [False, True, True])):
        op.set_preconditioner_threshold(thr)
        phi = torch.zeros_like(qm)
        st = op.multigroup_solve_dev(S, phi, qmom=qm, kind="fgmres", rtol=1e-9)
        assert all(s["status"] == 0 for s in st)
        assert [s["fallback"] != -1 for s in st] == used, thr
        for g in range(G):
            if not used[g]:
                assert torch.equal(phi[g], plain[g]), (thr, g)
            else:
                assert float((phi[g] - plain[g]).abs().max() / plain[g].abs().max()) <= 1e-6
    with pytest.raises(_lib.LionboltError):
        op.set_preconditioner_threshold(-1.0)
    op.close()
