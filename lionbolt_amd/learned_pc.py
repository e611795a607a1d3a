"""Learned right preconditioner for flux-moment flexible GMRES.

The network (18 -> 64 -> 64 -> 64 -> 16, tanh) is applied per spatial dof by the
library (``lb_mlp_apply_dev``): inputs are a residual's 16 flux moments at the
dof divided by the residual's RMS, log(sigma_t h) and sigma_s0/sigma_t of the
element; outputs are the error's 16 moments (times the RMS).

Training data come from the solver itself: flux-moment GMRES(m) runs on small
seeded problems and records its iterate x_k after every Arnoldi step
(``lb_gmres_snapshots_dev``).  With the converged solution phi*, every iterate
gives a pair residual r_k = b - A x_k (the library's operator, FP64) -> error
e_k = phi* - x_k.  Rows are the per-dof features of r_k (``lb_mlp_features_dev``)
and e_k / rms(r_k), subsampled per iterate.  Training is FP32 minibatch Adam in
the library's HIP kernels (``lb_mlp_train_dev``).  Everything here only moves
tensors between those calls.
"""
import math
import time

import numpy as np

from . import _lib, angular, problems, rng

NIN, NH, NOUT = 18, 64, 16
SHAPES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
          ("W4", (NH, NOUT)), ("b4", (NOUT,))]


def init_params(seed=2100):
    """Glorot-uniform weights (SplitMix64 stream ``seed``), zero biases."""
    out, k = [], 0
    for name, shp in SHAPES:
        size = int(np.prod(shp))
        if name.startswith("W"):
            lim = math.sqrt(6.0 / (shp[0] + shp[1]))
            out.append((2.0 * rng.uniform(seed, size, start=k) - 1.0) * lim)
            k += size
        else:
            out.append(np.zeros(size))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == _lib.MLP_PARAMS
    return p


def unpack(params):
    d, o = {}, 0
    for name, shp in SHAPES:
        size = int(np.prod(shp))
        d[name] = np.asarray(params[o:o + size]).reshape(shp)
        o += size
    return d


def forward_reference(params, X):
    """FP64 numpy forward pass (for tests)."""
    p = unpack(params)
    h = np.tanh(X @ p["W1"] + p["b1"])
    h = np.tanh(h @ p["W2"] + p["b2"])
    h = np.tanh(h @ p["W3"] + p["b3"])
    return h @ p["W4"] + p["b4"]


def _source_moments(op, m, ang, sigt, sigs, kind):
    """Single-group source moments (nmom * NENK,) of a training problem (cuda):
    the +z beam's first-collision source, or psi = 1 in every direction."""
    import torch
    if kind == "beam":
        S = np.zeros((1, 1) + sigs.shape)
        S[0, 0] = sigs
        q, _, _ = problems.beam_group_sources(op, m, S, np.atleast_2d(sigt))
        return q[0].reshape(-1).contiguous()
    q = torch.zeros((ang.n_moments, m.NENK), dtype=torch.float64, device="cuda")
    q[0] = math.sqrt(4.0 * math.pi)
    return q.reshape(-1)


def collect(op, qmom, m=30, rtol=1e-10, max_snaps=80, rows_per_snap=512, seed=0):
    """(features, targets) rows from one flux-moment GMRES solve on ``op``."""
    import torch
    ang_n = op.NI
    nphi = qmom.numel()
    q = torch.empty(op.NENK * ang_n, dtype=torch.float64, device="cuda")
    op.discrete_source_dev(qmom, q)
    phi = torch.empty(nphi, dtype=torch.float64, device="cuda")
    snaps = torch.empty((max_snaps, nphi), dtype=torch.float64, device="cuda")
    cnt, st = op.gmres_snapshots_dev(q, phi, snaps, m=m, rtol=rtol)
    b = torch.empty(nphi, dtype=torch.float64, device="cuda")
    op.moment_rhs_dev(q, b)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    Xs, Ys = [], []
    r = torch.empty_like(b)
    for k in range(cnt):
        x = snaps[k]
        op.moment_matvec_dev(x, r)
        r = b - r
        feat, s = op.mlp_features_dev(r)
        if not s > 0:
            continue
        e = ((phi - x) / s).view(16, -1).t()
        rows = torch.randint(0, op.NENK, (min(rows_per_snap, op.NENK),), generator=gen, device="cuda")
        Xs.append(feat[rows])
        Ys.append(e[rows].to(torch.float32))
    return Xs, Ys, st, cnt


def build_training_set(nmesh=32, seed=2000, rows_per_snap=512, L=3, log=None):
    """Runs the training problems and returns (X, Y) cuda float32 tensors."""
    import torch
    from .ops import Operator
    ang = problems.product_8x32() if L == 3 else angular.product(8, 32, L)
    Xs, Ys = [], []
    info = []
    t0 = time.perf_counter()
    for p, pr in enumerate(problems.pc_training_problems(nmesh, seed, L=L)):
        m = pr["mesh"]
        op = Operator(m, ang)
        op.set_physics(pr["sigt"], pr["sigs"])
        for j, kind in enumerate(("beam", "uniform")):
            qm = _source_moments(op, m, ang, pr["sigt"], pr["sigs"], kind)
            X, Y, st, cnt = collect(op, qm, rows_per_snap=rows_per_snap, seed=seed + 2 * p + j)
            Xs += X
            Ys += Y
            info.append({"problem": p, "source": kind, "tets": m.NE, "iterates": st["iterates"], "snapshots": cnt})
        op.close()
        if log:
            log(f"training problem {p}: {m.NE} tets, {info[-2]['iterates']}/{info[-1]['iterates']} iterates")
    X = torch.cat(Xs).contiguous()
    Y = torch.cat(Ys).contiguous()
    return X, Y, {"rows": int(X.shape[0]), "solves": info, "seconds": time.perf_counter() - t0}


def select_threshold(params, nmesh=32, seed=2000, L=3, fmt="bf16", rtol=1e-6):
    """The a-priori choice between the network and plain GMRES, from the training
    problems (beam sources): each is solved with flux-moment GMRES(30) and with
    FGMRES(30) and the network, both timed on the device after a warm-up solve.
    The threshold is the smallest optical thickness t (largest sigma_t times the
    cube root of the mesh volume, as the library measures a group) such that
    FGMRES was faster on every training problem at least t thick; inf (the network
    never used) when it was slower on the thickest.  Returns (threshold, table)."""
    import torch
    from .ops import Operator
    ang = problems.product_8x32() if L == 3 else angular.product(8, 32, L)
    table = []
    for p, pr in enumerate(problems.pc_training_problems(nmesh, seed, L=L)):
        m = pr["mesh"]
        op = Operator(m, ang)
        op.set_physics_groups(pr["sigt"][None, :], pr["sigs"][None, :, :])
        op.set_preconditioner(params, fmt)
        S = np.zeros((1, 1) + pr["sigs"].shape)
        S[0, 0] = pr["sigs"]
        qm = _source_moments(op, m, ang, pr["sigt"], pr["sigs"], "beam").reshape(1, op.nmom, op.NENK)
        secs, its = {}, {}
        for kind in ("gmres", "gmres", "fgmres"):             # the first solve settles the sweep autotuner
            phi = torch.zeros_like(qm)
            st = op.multigroup_solve_dev(S, phi, qmom=qm, kind=kind, space="moment", m=30, rtol=rtol, atol=0.0)
            secs[kind], its[kind] = st[0]["seconds"], st[0]["iterates"]
        tau = float(pr["sigt"].max() * np.cbrt(m.volumes().sum()))
        table.append({"problem": p, "thickness": tau, "gmres_s": secs["gmres"], "fgmres_s": secs["fgmres"],
                      "gmres_iterates": its["gmres"], "fgmres_iterates": its["fgmres"]})
        op.close()
    thr = math.inf
    for row in sorted(table, key=lambda r: -r["thickness"]):
        if row["fgmres_s"] < row["gmres_s"]:
            thr = row["thickness"]
        else:
            break
    return thr, table


def train(X, Y, steps=4000, batch=2048, lr=2e-3, seed=2100):
    """FP32 Adam in the library's kernels; targets are standardised by their RMS
    tau for training and tau is folded back into the output layer."""
    import torch
    tau = float(torch.sqrt((Y.double() ** 2).mean()))
    Yn = (Y / tau).contiguous()
    params = init_params(seed)
    loss = np.empty(steps, dtype=np.float32)
    t0 = time.perf_counter()
    _lib.check(_lib.lib().lb_mlp_train_dev(X.data_ptr(), Yn.data_ptr(), X.shape[0], params.ctypes.data_as(_lib._f),
                                           params.size, steps, batch, float(lr), seed, loss.ctypes.data_as(_lib._f)))
    secs = time.perf_counter() - t0
    o = sum(int(np.prod(s)) for _, s in SHAPES[:6])
    params[o:o + NH * NOUT] *= tau                  # W4
    params[o + NH * NOUT:] *= tau                   # b4
    return params, {"tau": tau, "loss_first": float(loss[:50].mean()), "loss_last": float(loss[-50:].mean()),
                    "steps": steps, "batch": batch, "seconds": secs}
