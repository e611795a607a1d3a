"""CPU baseline of the benchmark: the oracle (the C++/OpenMP FP64 restatement of
the reference's algorithm) timed on this host's cores.  Test infrastructure: run
by bench.py as a child process, so that its OpenMP settings (all cores of this
process's CPU set, OMP_PROC_BIND=spread) take effect before the runtime starts;
prints one JSON object.

Timed (steady clock, median of 3 after one warm-up):
* one transport sweep (TransportAlgebra.f90:29-106) of one group over all
  directions of the full mesh, on a vector already in sweep order;
* one dense scattering product (ScatteringAlgebra.f90:30-53, the reference's
  direction x direction kernel) and one moment-form product at full size;
* the 48,000-tet S8 problem solved end to end (psi-space GMRES(30), dense K);
* one sweep of each of the first four batch problems.
The schedule and geometry build is timed separately (the reference does it in
Build)."""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# run as a script, this file's own folder comes first on the path and would
# shadow the oracle package with its module of the same name
sys.path = [ROOT] + [p for p in sys.path if os.path.abspath(p or ".") != os.path.dirname(os.path.abspath(__file__))]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed(fn, reps=3):
    fn()                                   # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def main():
    cores = len(os.sched_getaffinity(0))   # before OpenMP binds this thread to one place
    from lionbolt_amd import angular, problems
    from oracle import oracle
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 40
    out = {"cpu_model": cpu_model(), "threads": oracle.num_threads(), "cores_available": cores,
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "")}
    m = problems.slab_cube_mesh(n)
    ang = problems.product_8x32()
    t0 = time.perf_counter()
    op = oracle.OracleOp(m, ang)
    out["build_s"] = time.perf_counter() - t0
    sigt, S = problems.downscatter_xs(G=1, L=ang.L)
    op.set_physics(sigt[0], S[0, 0])
    x = np.random.default_rng(1003).random(op.n)
    y = x.copy()

    def sweep():
        y[:] = x
        op.sweep_only(y)

    t_sweep, ts = timed(sweep)
    out["sweep"] = {"seconds": t_sweep, "runs": ts, "cell_angle_sweeps_per_s": m.NE * ang.NI / t_sweep,
                    "sample": f"1 group x {ang.NI} directions x {m.NE} tets (full sweep), median of 3"}
    t_k, ts = timed(lambda: op.scatter_matvec(x))
    out["dense_K"] = {"seconds": t_k, "runs": ts}
    D, Mm = angular.moment_matrices(ang)
    deg = angular.moment_degree(ang.L)
    t_km, ts = timed(lambda: op.scatter_moment_matvec(x, D, Mm, deg))
    out["moment_K"] = {"seconds": t_km, "runs": ts}
    del op, x, y
    m48, a48, st48, ss48, q48 = problems.cube_source_48k()
    o48 = oracle.OracleOp(m48, a48)
    o48.set_physics(st48, ss48)
    t0 = time.perf_counter()
    _, st = o48.boltzmann_matinv(q48, kind="gmres", m=30, rtol=1e-8)
    out["cube48k_solve"] = {"seconds": time.perf_counter() - t0, "iterates": st["iters"], "status": st["status"],
                            "method": "psi-space GMRES(30), dense K, tol 1e-8"}
    s8 = angular.level_symmetric_s8()
    tot_ca, tot_s = 0, 0.0
    for pr in problems.uneven_beams()[:4]:
        o = oracle.OracleOp(pr["mesh"], s8)
        o.set_physics([pr["sigt"]], [[pr["c"] * pr["sigt"]]])
        v = np.random.default_rng(1004).random(o.n)
        t0 = time.perf_counter()
        o.sweep_only(v)
        tot_s += time.perf_counter() - t0
        tot_ca += pr["mesh"].NE * s8.NI
    out["uneven_beams_sweep"] = {"cell_angle_sweeps_per_s": tot_ca / tot_s,
                                 "sample": "problems 1-4 of the 64, one sweep each, all 80 directions"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
