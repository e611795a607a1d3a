"""Throughput of the transport sweep on MI355X: cell-angle-group sweeps per second.

Workload (default): the 384,000-tet jittered cube (40^3 hexes x 6 Kuhn tets,
seed 1001, a 3x-density slab z in [4, 6] cm as material 2), 256 directions
(8 Gauss-Legendre polar x 32 Chebyshev azimuthal), 24 groups with sigma_t
log-uniform on [0.02, 0.5] /cm (seed 1002).  One step = one transport sweep
T^-1 of all 24 groups x 256 directions x 384,000 cells (2.36e9
cell-angle-groups) with the source and the result resident in HBM.

The line also carries:
* roofline: the sweep kernel's algorithmic HBM bytes (64 B per
  cell-angle-group: four FP64 source values in, four FP64 flux values out)
  over the sweep time measured with HIP events on the operator's stream,
  against the 8 TB/s HBM3E peak;
* cpu_baseline: the FP64 C++/OpenMP oracle (restatement of the reference's
  per-direction OpenMP sweep) timed on this host on a bounded sample: one
  group, 32 of the 256 directions, the full 384k-tet mesh;
* solve: time to a 1e-6 relative transport residual of the single-group
  within-group solve (plain GMRES on the device) on the same mesh.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CAG = 64


def cfg2_inputs(n=40, G=24):
    from lionbolt_amd import angular, mesh, rng
    m = mesh.jittered_cube(n, a=10.0, jitter=0.2, seed=1001, slab=(4.0, 6.0))
    ang = angular.product(8, 32, 3)
    u = rng.uniform(1002, G)
    sigt1 = np.exp(np.log(0.02) + u * (np.log(0.5) - np.log(0.02)))     # log-uniform [0.02, 0.5]
    sigt = np.stack([sigt1, 3.0 * sigt1], axis=1)                        # slab: 3x density
    return m, ang, sigt


def cpu_baseline(m, ang, ndirs=32, reps=2):
    from lionbolt_amd import angular
    from oracle import oracle
    sub = angular.AngularSet(ang.k[:: ang.NI // ndirs][:ndirs], ang.w[:: ang.NI // ndirs][:ndirs], 0)
    op = oracle.OracleOp(m, sub)
    op.set_physics([0.2, 0.6], [[0.0], [0.0]])
    q = np.random.default_rng(1003).random(op.n)
    op.sweep_only(q.copy())          # warm-up
    times = []
    for _ in range(reps):
        x = q.copy()
        t0 = time.perf_counter()
        op.sweep_only(x)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    cag = m.NE * sub.NI
    return {"value": cag / t, "unit": "cell-angle-group sweeps/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"1 group x {sub.NI} of 256 directions x {m.NE} tets, FP64 oracle sweep, median of {reps}"}


def run(args):
    import torch
    from lionbolt_amd.ops import Operator

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    m, ang, sigt = cfg2_inputs(args.n, args.groups)
    G = args.groups
    op = Operator(m, ang)
    sigs = np.zeros((G, 2, ang.L + 1))
    op.set_physics_groups(sigt, sigs)
    setup_s = time.perf_counter() - t0
    cag = m.NE * ang.NI * G
    nvals = cag * 4
    q = torch.rand(nvals, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(q)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        op.transport_matinv_dev(q, psi, 0, G)
    op.synchronize()
    torch.cuda.synchronize()
    sweep_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.transport_matinv_dev(q, psi, 0, G)
        sweep_ms.append(op.last_sweep_ms())     # HIP events around the level launches
    op.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1000.0 * wall / args.steps
    value = cag * args.steps / wall
    ev_ms = float(np.mean(sweep_ms))
    achieved = BYTES_PER_CAG * cag / (ev_ms * 1e-3) / 1e9
    if not torch.isfinite(psi).all():
        raise RuntimeError("non-finite flux")
    del q, psi
    torch.cuda.empty_cache()
    out = {
        "metric": "cell-angle-group sweeps/s",
        "value": value,
        "unit": "cell-angle-group sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered-cube mesh and cross sections)",
        "config": {"workload": f"DFEM SN sweep, {m.NE} tets, {ang.NI} directions (GL8 x C32), {G} groups, "
                               f"P{ang.L}, level-synchronous, in HBM",
                   "tets": m.NE, "directions": ang.NI, "groups": G, "levels_max": op.maxlev,
                   "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "sweep_level_kernel", "bytes_per_unit": BYTES_PER_CAG,
                     "event_ms_per_sweep": ev_ms},
    }
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(m, ang)
    op.close()
    print(json.dumps(out))


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=40, help="hexes per side (40 -> 384,000 tets)")
    p.add_argument("--groups", type=int, default=24)
    p.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    args = p.parse_args()
    if int(os.environ.get("WORLD_SIZE", "1")) > 1 or args.gpus > 1:
        print(json.dumps({"error": "one GPU per run: nothing here uses a second GPU"}))
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
