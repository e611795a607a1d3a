"""Batches of independent problems: the persistent work-queue sweep and the
lockstep per-problem GMRES, checked against the level launches of the same
union mesh (bit for bit) and against each problem solved alone."""
import math

import numpy as np
import pytest

from lionbolt_amd import _lib, angular, mesh as M, problems

pytestmark = pytest.mark.gpu


def small_batch():
    """Five uneven cubes with their own cross sections and beams."""
    specs = [(2, 3.0, 0.9, 0.5), (4, 6.0, 0.3, 0.9), (3, 2.0, 1.5, 0.7), (5, 8.0, 0.1, 0.95), (3, 4.0, 0.6, 0.6)]
    probs = []
    for p, (n, a, st, c) in enumerate(specs):
        fi = 0.7 + 1.1 * p
        mu = 0.55 + 0.08 * p
        s = math.sqrt(1 - mu * mu)
        probs.append({"n": n, "a": a, "sigt": st, "c": c, "k0": np.array([s * math.cos(fi), s * math.sin(fi), mu]),
                      "mesh": M.jittered_cube(n, a=a, seed=3000 + p)})
    return problems.uneven_beams_batch(probs=probs)


def batch_op(um, eoff, sigt, sigs, ang):
    from lionbolt_amd.ops import Operator
    op = Operator(um, ang)
    op.set_physics(sigt, sigs)
    op.set_problems(eoff)
    return op


@pytest.mark.parametrize("block,per_cu", [(256, 4), (256, 8), (512, 2), (1024, 1), (1024, 2)])
def test_persistent_sweep_equals_level_launches(block, per_cu):
    import torch
    um, eoff, sigt, sigs, probs = small_batch()
    ang = angular.level_symmetric_s8()
    op = batch_op(um, eoff, sigt, sigs, ang)
    q = torch.rand(op.n, dtype=torch.float64, device="cuda")
    op.set_sweep_mode("levels")
    ref = torch.empty_like(q)
    op.transport_matinv_dev(q, ref, 0, 1)
    op.synchronize()
    ref = ref.cpu().numpy()
    op.set_direction_block(block)
    op.set_workgroups_per_cu(per_cu)
    op.set_sweep_mode("persistent")
    for _ in range(3):                        # graph replays with eager work between
        psi = torch.full_like(q, float("nan"))
        op.transport_matinv_dev(q, psi, 0, 1)
        op.synchronize()
        assert np.array_equal(psi.cpu().numpy(), ref)


def test_batch_sweep_equals_each_problem_alone():
    import torch
    from lionbolt_amd.ops import Operator
    um, eoff, sigt, sigs, probs = small_batch()
    ang = angular.level_symmetric_s8()
    op = batch_op(um, eoff, sigt, sigs, ang)
    op.set_sweep_mode("auto")
    q = np.random.default_rng(5).random(op.n)
    qd = torch.from_numpy(q).cuda()
    for _ in range(20):                       # through the autotuner's timing phase and past it
        psi = torch.empty_like(qd)
        op.transport_matinv_dev(qd, psi, 0, 1)
        op.synchronize()
    mode, threads, per_cu = op.sweep_choice(1)
    assert mode in ("graph", "persistent", "direction")
    batch = psi.cpu().numpy().reshape(ang.NI, um.NENK)
    for p, pr in enumerate(probs):
        m = pr["mesh"]
        single = Operator(m, ang)
        single.set_physics([pr["sigt"]], [[pr["c"] * pr["sigt"]]])
        cols = slice(4 * eoff[p], 4 * eoff[p + 1])
        qp = np.ascontiguousarray(q.reshape(ang.NI, um.NENK)[:, cols]).reshape(-1)
        assert np.array_equal(single.transport_matinv(qp).reshape(ang.NI, -1), batch[:, cols])
        single.close()


def test_batch_solve_matches_single_problem_solves():
    import torch
    from lionbolt_amd.ops import Operator
    um, eoff, sigt, sigs, probs = small_batch()
    ang = angular.level_symmetric_s8()
    op = batch_op(um, eoff, sigt, sigs, ang)
    qd = problems.uneven_beams_sources(op, eoff, probs, sigt, sigs).reshape(-1).contiguous()
    qmom = qd.cpu().numpy()[None, :]
    phi = torch.zeros_like(qd)
    stats = op.batch_solve_dev(qd, phi, m=30, rtol=1e-10)
    assert all(s["status"] == 0 for s in stats)
    assert max(s["true_residual"] for s in stats) <= 1e-8
    scalar = phi.cpu().numpy() * math.sqrt(4 * math.pi)      # phi_00 = Y00 sum_i w_i psi_i
    y00 = 1.0 / math.sqrt(4 * math.pi)
    for p, pr in enumerate(probs):
        m = pr["mesh"]
        single = Operator(m, ang)
        single.set_physics([pr["sigt"]], [[pr["c"] * pr["sigt"]]])
        qsrc = np.tile(y00 * qmom[0, 4 * eoff[p]:4 * eoff[p + 1]], ang.NI)
        psi, st = single.boltzmann_matinv(qsrc, kind="gmres", space="moment", m=30, rtol=1e-10)
        ref = single.fluence(psi)
        got = scalar[4 * eoff[p]:4 * eoff[p + 1]]
        assert np.abs(got - ref).max() <= 1e-7 * np.abs(ref).max(), p
        single.close()


def test_problem_set_validation():
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(3, a=2.0)
    ang = angular.snprep(1)
    op = Operator(m, ang)
    op.set_physics([1.0], [[0.5, 0.0]])
    with pytest.raises(_lib.LionboltError) as ei:
        op.set_problems([0, 50, m.NE])        # a face joins the two ranges
    assert ei.value.code == -1
    with pytest.raises(_lib.LionboltError):
        op.set_problems([0, m.NE - 1])        # does not end at NE
    op.set_sweep_mode("persistent")
    import torch
    q = torch.zeros(op.n, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.LionboltError):
        op.transport_matinv_dev(q, q, 0, 1)   # no problem set
