"""Every way of issuing a sweep produces the same flux, bit for bit: one kernel
per global level (streamed or replayed from a captured graph) and the
direction-resident kernels in which a workgroup walks all levels of one
direction with workgroup barriers, reading upstream traces from global memory
or from the LDS slots of the slot plan (ring slots, staging copies, and global
reads where neither holds the trace)."""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("G,block", [(1, 256), (1, 512), (1, 768), (1, 1024), (2, 512), (3, 768), (5, 512),
                                     (24, 1024), (24, 768), (24, 256)])
def test_direction_resident_equals_level_launches(G, block):
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(6, a=5.0, seed=1007, slab=(2.0, 3.0))
    ang = angular.product(4, 8, 1)
    op = Operator(m, ang)
    sigt = np.stack([np.linspace(0.1, 1.5, G), np.linspace(0.3, 4.5, G)], axis=1)
    op.set_physics_groups(sigt, np.zeros((G, 2, ang.L + 1)))
    op.set_direction_block(block)
    q = torch.rand(ang.NI * m.NE * 4 * G, dtype=torch.float64, device="cuda")
    res = {}
    modes = ["levels", "direction", "direction", "graph", "direction_lds", "direction_lds"]
    for mode in modes:
        op.set_sweep_mode(mode)
        psi = torch.full_like(q, float("nan"))
        op.transport_matinv_dev(q, psi, 0, G)
        op.synchronize()
        res.setdefault(mode, []).append(psi.cpu().numpy())
    for mode in set(modes) - {"levels"}:
        for v in res[mode]:
            assert np.array_equal(v, res["levels"][0]), mode


@pytest.mark.parametrize("G", [1, 3])
def test_lds_slots_with_staging_and_global_reads(G):
    """82,944 tets, 8 directions.  With one group every trace fits the ring; with
    three the ring holds 1,664 cells, so the plan also evicts, stages copies from
    global memory and leaves some traces in global memory (checked on the host
    plan, whose counts do not depend on the order of cells inside a level)."""
    import torch
    from lionbolt_amd.ops import Operator, trace_plan_host
    m = M.jittered_cube(24, a=6.0, seed=1013, slab=(2.0, 3.5))
    ang = angular.product(2, 4, 0)
    stats = trace_plan_host(m, ang, G, 160 * 1024)["stats"]
    if G == 3:
        assert stats[1] > 0 and stats[2] > 0 and stats[3] > 0, stats
    op = Operator(m, ang)
    sigt = np.stack([np.linspace(0.3, 0.9, G), np.linspace(0.8, 2.5, G)], axis=1)
    op.set_physics_groups(sigt, np.zeros((G, 2, 1)))
    q = torch.rand(ang.NI * m.NE * 4 * G, dtype=torch.float64, device="cuda")
    out = {}
    for mode in ("levels", "direction_lds"):
        op.set_sweep_mode(mode)
        for block in (256, 512, 768):
            op.set_direction_block(block)
            psi = torch.full_like(q, float("nan"))
            op.transport_matinv_dev(q, psi, 0, G)
            op.synchronize()
            out[(mode, block)] = psi.cpu().numpy()
    for block in (256, 512, 768):
        assert np.array_equal(out[("direction_lds", block)], out[("levels", 256)])


def test_direction_resident_in_place_and_solver():
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(5, a=4.0)
    ang = angular.snprep(3)
    op = Operator(m, ang)
    op.set_physics([1.0], [[0.8, 0.3, 0.1, 0.05]])
    q = np.random.default_rng(3).random(op.n)
    op.set_sweep_mode("levels")
    a, sa = op.boltzmann_matinv(q, space="moment", m=30)
    t1 = op.transport_matinv(q)
    for mode in ("direction", "direction_lds"):
        op.set_sweep_mode(mode)
        b, sb = op.boltzmann_matinv(q, space="moment", m=30)
        t2 = op.transport_matinv(q)             # in place on the staging buffer
        assert sa["iterates"] == sb["iterates"]
        assert np.array_equal(a, b)
        assert np.array_equal(t1, t2)


@pytest.mark.parametrize("G", [1, 24])
def test_autotuned_sweeps_equal_level_launches(G):
    """The autotuner times every configuration on live sweeps; every sweep it
    issues (timed or settled) matches the level launches bit for bit."""
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(5, a=4.0, seed=1011, slab=(1.5, 2.5))
    ang = angular.product(4, 8, 1)
    op = Operator(m, ang)
    sigt = np.stack([np.linspace(0.2, 1.1, G), np.linspace(0.5, 3.5, G)], axis=1)
    op.set_physics_groups(sigt, np.zeros((G, 2, ang.L + 1)))
    q = torch.rand(ang.NI * m.NE * 4 * G, dtype=torch.float64, device="cuda")
    op.set_sweep_mode("levels")
    ref = torch.empty_like(q)
    op.transport_matinv_dev(q, ref, 0, G)
    op.synchronize()
    ref = ref.cpu().numpy()
    op.set_sweep_mode("auto")
    assert op.sweep_choice(G) == ("auto", 0, 0)
    for _ in range(16):     # two timed sweeps for each of the seven configurations, then settled ones
        psi = torch.full_like(q, float("nan"))
        op.transport_matinv_dev(q, psi, 0, G)
        op.synchronize()
        assert np.array_equal(psi.cpu().numpy(), ref)
    mode, threads, _ = op.sweep_choice(G)
    assert (mode, threads) in {("graph", 0), ("direction", 256), ("direction", 512), ("direction", 1024),
                               ("direction_lds", 256), ("direction_lds", 512), ("direction_lds", 768)}
