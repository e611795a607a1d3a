"""Multigroup downscatter solve on the GPU against the oracle driven like the
reference's group loop (Main.f90:326-476): for each group g the source is the
external source plus sum_{g'<g} K_{g'->g} psi_{g'} applied with the dense
transfer kernel (PIEISource, Main.f90:584-629), then BoltzmannOp%MatInv."""
import numpy as np
import pytest

from lionbolt_amd import angular, mesh as M, problems
from oracle.oracle import OracleOp

pytestmark = pytest.mark.gpu


def beam_qmom(m, ang, S, sigt):
    """The +z beam's group source moments (problems.beam_group_sources) as a host array."""
    from lionbolt_amd.ops import Operator
    op = Operator(m, ang)
    q = problems.beam_group_sources(op, m, S, sigt)[0].cpu().numpy()
    op.close()
    return q


def oracle_multigroup(m, ang, sigt, S, qmom, rtol):
    G = S.shape[0]
    _, Mm = angular.moment_matrices(ang)
    within = OracleOp(m, ang)
    coupling = OracleOp(m, ang)
    psis = []
    for g in range(G):
        s = (Mm @ qmom[g]).reshape(-1)
        for gp in range(g):
            coupling.set_physics(np.ones(sigt.shape[1]), S[gp, g])
            s += coupling.scatter_matvec(psis[gp])
        within.set_physics(sigt[g], S[g, g])
        x, st = within.boltzmann_matinv(s, kind="gmres", m=30, rtol=rtol)
        assert st["status"] == 0
        psis.append(x)
    return psis


def test_multigroup_downscatter_matches_oracle():
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(4, a=10.0, seed=1001, slab=(4.0, 6.0))
    ang = angular.product(2, 8, 1)
    G = 4
    sigt, S = problems.downscatter_xs(G=G, L=ang.L)
    sigt = sigt * 10.0                   # optically thicker so groups couple noticeably
    S = S * 10.0
    qmom = beam_qmom(m, ang, S, sigt)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, problems.within_group(S))
    phi = torch.zeros((G, ang.n_moments, m.NENK), dtype=torch.float64, device="cuda")
    q_d = torch.from_numpy(qmom).cuda()
    stats = op.multigroup_solve_dev(S, phi, qmom=q_d, rtol=1e-11)
    assert all(s["status"] == 0 for s in stats)
    psis = oracle_multigroup(m, ang, sigt, S, qmom, rtol=1e-11)
    D, _ = angular.moment_matrices(ang)
    got = phi.cpu().numpy()
    for g in range(G):
        ref = D @ psis[g].reshape(ang.NI, -1)
        assert np.abs(got[g] - ref).max() <= 1e-7 * np.abs(ref).max(), g
        assert stats[g]["true_residual"] <= 1e-10


@pytest.mark.parametrize("batch", [2, 3, 5])
def test_group_batched_solve_matches_group_by_group(batch):
    """Batches of consecutive groups solved as one coupled system reach the same
    fluxes as the group-by-group solve, each group to its own tolerance."""
    import torch
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(4, a=10.0, seed=1001, slab=(4.0, 6.0))
    ang = angular.product(2, 8, 3)
    G = 5
    sigt, S = problems.downscatter_xs(G=G, L=ang.L)
    sigt = sigt * 10.0
    S = S * 10.0
    qmom = beam_qmom(m, ang, S, sigt)
    q_d = torch.from_numpy(qmom).cuda()
    res = {}
    for b in (1, batch):
        op = Operator(m, ang)
        op.set_physics_groups(sigt, problems.within_group(S))
        op.set_group_batch(b)
        phi = torch.zeros((G, ang.n_moments, m.NENK), dtype=torch.float64, device="cuda")
        stats = op.multigroup_solve_dev(S, phi, qmom=q_d, rtol=1e-10, atol=0.0)
        assert all(s["status"] == 0 for s in stats)
        assert max(s["true_residual"] for s in stats) <= 1e-9
        res[b] = phi.cpu().numpy()
        op.close()
    for g in range(G):
        ref = res[1][g]
        assert np.abs(res[batch][g] - ref).max() <= 1e-8 * np.abs(ref).max(), g
