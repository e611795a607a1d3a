"""Seeded synthetic workloads of the transport-solve path.

Decisions recorded here (the reference never ran these configurations):

* cube_source_48k: 10 cm cube, 20^3 hexes x 6 tets (48,000), LQ S8, one group, P0,
  sigma_t = 1.0, sigma_s = 0.9 /cm, vacuum; source q_i = 1/(4 pi) on every node
  of elements whose centroid lies in the central cube [4, 6]^3.
* slab_cube_mesh / product_8x32 / downscatter_xs: 40^3 x 6 = 384,000 tets, Gauss-Legendre 8 polar x Chebyshev 32 azimuthal
  (256 directions), P3, 24 groups, downscatter only.  sigma_t,g log-uniform on
  [0.02, 0.5] /cm (SplitMix64 stream 1002); the slab z in [4, 6] cm (elements by
  centroid) is material 2 with 3x all cross sections.  Transfer moments
  sigma_0(g' -> g) for g >= g' follow 0.5^(g - g') normalised so that the
  out-scatter of g' is 0.85 sigma_t,g' (the "row sum" read as out-scatter);
  sigma_l = 0.6^l sigma_0.
* first-collision beam: a unit mono-directional planar beam along +z entering
  the face z = 0 in group 1 (source plane z = -1, no cutout), built as the
  reference builds it (beam.py over lb_beam_*: the uncollided fluence ray-traced
  to the tetrahedral quadrature points and projected with IMr,
  FCSSource.f90:134-176).  Group 1's source is its first-collision source
  sigma_l(1 -> 1) Y_lm(z) phi_0 (FCSSource.f90:178-215); group g > 1 receives
  sigma_l(1 -> g) times the moments of group 1's uncollided angular fluence, which
  the reference adds to the solution (Main.f90:443) before the energy coupling
  applies K(1 -> g) to it (Main.f90:584-629).
* uneven_beams: 64 single-group problems (SplitMix64 stream 1004, six values per
  problem): tets drawn uniform on [2,000, 60,000] and quantised to the jittered
  cube 6 n^3 with n = round((N/6)^(1/3)) clipped to [7, 21]; edge length
  U[5, 20] cm; sigma_t U[0.05, 1.0] /cm; c = sigma_s/sigma_t U[0.5, 0.95]; P0;
  LQ S8; a broad collimated planar beam of unit flux along
  k0 = (sin t cos f, sin t sin f, cos t) with cos t U[0.5, 1.0] and f U[0, 2 pi)
  entering every face it illuminates (no cutout, source plane 2a upstream of the
  cube's centre); jitter stream 1001 + 7919 (p + 1).  Its first-collision source
  is built as the reference builds it (beam.py); with P0 scattering it is
  isotropic, sigma_s Y_00 phi_0 in moments.
* scattering_dominated_xs: the 384,000-tet mesh, directions, groups and sigma_t with a within-group
  scattering ratio sigma_s0(g -> g) / sigma_t,g = 0.99 in every group (the
  scattering-dominated regime the configuration asks for) and a weak downscatter
  tail of 0.005 sigma_t,g' (shares 0.5^(g - g') as downscatter_xs) so that the +z beam's
  first-collision source reaches every group; sigma_l = 0.6^l sigma_0.
* learned-preconditioner training problems: 32 jittered cubes (10 cm, n = 10..16
  hexes per side, jitter stream 2000 + p, the 3x slab) with one group each,
  sigma_t log-uniform on [0.02, 0.5] /cm and the scattering-dominated ratios (SplitMix64 stream
  2000); sources: the +z beam's first-collision source and a uniform isotropic
  source.
"""
import math

import numpy as np

from . import angular, mesh, rng


def cube_source_48k():
    m = mesh.jittered_cube(20, a=10.0, jitter=0.2, seed=1001)
    ang = angular.level_symmetric_s8(0)
    c = m.centroids()
    inside = np.all((c >= 4.0) & (c <= 6.0), axis=1)
    q = np.tile(np.repeat(np.where(inside, 1.0 / (4 * math.pi), 0.0), 4), ang.NI)
    return m, ang, np.array([1.0]), np.array([[0.9]]), q


def slab_cube_mesh(n=40):
    return mesh.jittered_cube(n, a=10.0, jitter=0.2, seed=1001, slab=(4.0, 6.0))


def product_8x32():
    return angular.product(8, 32, 3)


def downscatter_xs(G=24, L=3, ratio=0.85, within_only=False):
    """sigma_t (G, 2) and transfer moments S (G, G, 2, L+1) indexed [g', g, mat, l]."""
    u = rng.uniform(1002, G)
    st = np.exp(np.log(0.02) + u * (np.log(0.5) - np.log(0.02)))
    sigt = np.stack([st, 3.0 * st], axis=1)
    S = np.zeros((G, G, 2, L + 1))
    for gp in range(G):
        if within_only:
            share = np.zeros(G)
            share[gp] = 1.0
        else:
            w = 0.5 ** np.arange(G - gp)
            share = np.zeros(G)
            share[gp:] = w / w.sum()
        for g in range(gp, G):
            for mat in range(2):
                s0 = ratio * sigt[gp, mat] * share[g]
                S[gp, g, mat, :] = s0 * 0.6 ** np.arange(L + 1)
    return sigt, S


def within_group(S):
    """Within-group moments (G, Nm, L+1) from a transfer table."""
    G = S.shape[0]
    return np.stack([S[g, g] for g in range(G)])


def plus_z_beam(a=10.0):
    """The unit planar beam along +z entering the face z = 0 of the cube [0, a]^3."""
    from .beam import ExternalBeam, FieldGeometry
    return ExternalBeam(FieldGeometry([0.0, 0.0, 1.0], [0.5 * a, 0.5 * a, -1.0]), n=1.0)


def beam_group_sources(op, m, S, sigt, beam=None):
    """Source moments (G, nmom, NENK) of a beam in group 1 (cuda float64), with its
    projected uncollided fluence phi0 (NENK) and the moments of its uncollided
    angular fluence (nmom, NENK): group 1 gets its first-collision source, group
    g > 1 the scattering of the uncollided fluence into g (see the module docstring).
    ``op`` is any operator on ``m`` with the quadrature of the solve."""
    import torch
    from . import angular
    beam = plus_z_beam() if beam is None else beam
    phi0 = beam.uncollided(op, sigt[0])
    fcs, umom = beam.moments(op, phi0)
    G = S.shape[0]
    deg = torch.from_numpy(angular.moment_degree(op.ang.L)).cuda()
    mat_dof = torch.from_numpy(np.repeat(m.mat - 1, 4)).cuda()
    S0 = torch.from_numpy(np.ascontiguousarray(S[0])).cuda()          # (G, Nm, L+1) for 1 -> g
    q = torch.empty((G, op.nmom, m.NENK), dtype=torch.float64, device="cuda")
    for g in range(G):
        q[g] = S0[g][mat_dof][:, deg].T * (fcs if g == 0 else umom)
    return q, phi0, umom


def uneven_beams(nprob=64, seed=1004):
    """The 64 uneven single-group beam problems (see the module docstring)."""
    u = rng.uniform(seed, 6 * nprob).reshape(nprob, 6)
    out = []
    for p in range(nprob):
        target = 2000.0 + 58000.0 * u[p, 0]
        n = int(np.clip(round((target / 6.0) ** (1.0 / 3.0)), 7, 21))
        a = 5.0 + 15.0 * u[p, 1]
        sigt = 0.05 + 0.95 * u[p, 2]
        c = 0.5 + 0.45 * u[p, 3]
        mu = 0.5 + 0.5 * u[p, 4]
        fi = 2.0 * math.pi * u[p, 5]
        st = math.sqrt(1.0 - mu * mu)
        k0 = np.array([st * math.cos(fi), st * math.sin(fi), mu])
        out.append({"n": n, "a": a, "sigt": sigt, "c": c, "k0": k0,
                    "mesh": mesh.jittered_cube(n, a=a, jitter=0.2, seed=1001 + 7919 * (p + 1))})
    return out


def uneven_beams_batch(nprob=64, seed=1004, probs=None):
    """The uneven-beam problems as one batch: union mesh, element offsets, per-problem
    (material) sigma_t and P0 sigma_s, and the problems."""
    probs = uneven_beams(nprob, seed) if probs is None else probs
    um, eoff = mesh.union([p["mesh"] for p in probs])
    sigt = np.array([p["sigt"] for p in probs])
    sigs = np.array([[p["c"] * p["sigt"]] for p in probs])
    return um, eoff, sigt, sigs, probs


def uneven_beams_sources(op, eoff, probs, sigt, sigs):
    """First-collision source moments (1, NENK) of the batch (cuda): each problem's
    own beam traced through the union mesh (rays never leave a problem: no face
    joins two), its projected fluence kept on that problem's elements; P0, so the
    source is sigma_s Y_00 phi_0."""
    import torch
    from .beam import ExternalBeam, FieldGeometry
    phi0 = torch.zeros(op.NENK, dtype=torch.float64, device="cuda")
    for p, pr in enumerate(probs):
        c = 0.5 * pr["a"] * np.ones(3)
        beam = ExternalBeam(FieldGeometry(pr["k0"], c - 2.0 * pr["a"] * pr["k0"]), n=1.0)
        beam.uncollided(op, sigt, elems=(int(eoff[p]), int(eoff[p + 1])), out=phi0)
    sig_dof = torch.from_numpy(np.repeat(sigs[:, 0][np.asarray(op.mesh.mat) - 1], 4)).cuda()
    return ((1.0 / math.sqrt(4.0 * math.pi)) * sig_dof * phi0)[None, :]


def scattering_dominated_xs(G=24, L=3, within=0.99, down=0.005):
    """Scattering-dominated cross sections: sigma_t (G, 2) of downscatter_xs and transfer moments (G, G, 2, L+1)."""
    sigt, _ = downscatter_xs(G=G, L=L)
    S = np.zeros((G, G, 2, L + 1))
    for gp in range(G):
        w = 0.5 ** np.arange(1, G - gp)
        for mat in range(2):
            S[gp, gp, mat, :] = within * sigt[gp, mat] * 0.6 ** np.arange(L + 1)
            if gp + 1 < G:
                for g in range(gp + 1, G):
                    S[gp, g, mat, :] = down * sigt[gp, mat] * (w[g - gp - 1] / w.sum()) * 0.6 ** np.arange(L + 1)
    return sigt, S


def pc_training_problems(nmesh=32, seed=2000, L=3, within=0.99):
    """Small single-group problems for training the learned preconditioner."""
    u = rng.uniform(seed, 2 * nmesh).reshape(nmesh, 2)
    out = []
    for p in range(nmesh):
        n = 10 + int(u[p, 0] * 7)
        st = float(np.exp(np.log(0.02) + u[p, 1] * (np.log(0.5) - np.log(0.02))))
        sigt = np.array([st, 3.0 * st])
        sigs = np.stack([within * sigt[0] * 0.6 ** np.arange(L + 1), within * sigt[1] * 0.6 ** np.arange(L + 1)])
        out.append({"n": n, "sigt": sigt, "sigs": sigs,
                    "mesh": mesh.jittered_cube(n, a=10.0, jitter=0.2, seed=seed + p, slab=(4.0, 6.0))})
    return out
