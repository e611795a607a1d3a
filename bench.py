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
  per-direction OpenMP sweep) timed on this host's cores (child process,
  OMP_PROC_BIND=spread, CPU model recorded): one full sweep of one group over
  the 384k-tet mesh and all 256 directions, and the full-size dense and
  moment-form scattering products, median of 3 after a warm-up;
* solve: time to a 1e-6 relative transport residual of the full 24-group,
  P3, downscatter problem with the first-collision source of a +z beam built
  as the reference builds it (group-by-group within-group GMRES on flux
  moments, plain, FP64), with the collided and uncollided scalar flux sums;
* reference_method_gpu: the same problem with the reference's own method
  (psi-space GMRES(10) per group), for its iteration counts;
* cube48k_gpu: the 48,000-tet S8 problem solved end to end (GMRES(30), 1e-8);
* learned_pc: the 24-group problem at within-group c = 0.99 solved with
  plain flux-moment GMRES(30) and with flexible GMRES(30) preconditioned by the
  per-cell-node network (trained first on the solver's residual/error pairs
  from 32 smaller problems; training time reported separately);
* uneven_beams_batch: 64 uneven single-group S8 beam problems (2k-56k tets) swept as
  one batch (persistent work-queue kernel, autotuned) and solved with
  per-problem GMRES in lockstep;
* high_order_operator: one sweep and the two moment products at the
  scattering orders of the reference's input decks (7, 11, 15) on an
  82,944-tet cube;
* cpu_baseline also carries the oracle's 48,000-tet solve time, projected
  (lower-bound) CPU times of the 24-group solve by the reference's method and
  by the GPU solve's own algorithm, and the oracle's sweep rate on the batch
  problems.
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


def cpu_baseline(n, ref_method):
    """The oracle (FP64 C++/OpenMP restatement of the reference's algorithm) on this
    host, run as a child process (oracle/cpu_baseline.py) with OMP_PROC_BIND=spread
    and the thread count the environment gives: the full one-group sweep of the
    384k-tet mesh over all 256 directions and the full-size dense and moment-form
    scattering products, median of 3 after a warm-up; the 48,000-tet problem
    solved end to end; the batch problems' sweep rate.  The 24-group solve is
    projected from these: the reference's own method (psi-space GMRES(10), dense
    K) at the iteration counts that method takes on the GPU in this run, and the
    GPU solve's algorithm (moment-form K) at its sweep count; both count sweeps and
    K applications only (no Krylov vector work), so they are lower bounds."""
    import subprocess
    env = dict(os.environ)
    env["OMP_PROC_BIND"] = "spread"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), str(n)], env=env,
                       capture_output=True, text=True, timeout=1200)
    if r.returncode != 0:
        return {"error": r.stderr[-2000:]}
    c = json.loads(r.stdout.strip().splitlines()[-1])
    sw = c["sweep"]
    out = {"value": sw["cell_angle_sweeps_per_s"], "unit": "cell-angle-group sweeps/s", "cores": c["threads"],
           "kind": "port", "sample": sw["sample"], "cpu_model": c["cpu_model"],
           "cores_available": c["cores_available"], "omp_proc_bind": c["omp_proc_bind"],
           "build_s": c["build_s"], "sweep_s": sw["seconds"], "sweep_runs_s": sw["runs"],
           "dense_K_s": c["dense_K"]["seconds"], "moment_K_s": c["moment_K"]["seconds"],
           "cube48k_solve": c["cube48k_solve"], "uneven_beams_sweep": c["uneven_beams_sweep"]}
    if ref_method:
        sweeps = ref_method["sweeps"]
        out["solve_projected"] = {
            "seconds": sweeps * (sw["seconds"] + c["dense_K"]["seconds"]), "sweeps_and_K": sweeps,
            "cpu_sweep_s": sw["seconds"], "cpu_dense_K_s": c["dense_K"]["seconds"],
            "note": "lower bound: full-size sweep + dense K times at the reference method's measured iteration "
                    "counts"}
        if ref_method.get("plain_sweeps"):
            n_pl = ref_method["plain_sweeps"]
            out["solve_projected_moment_form"] = {
                "seconds": n_pl * (sw["seconds"] + c["moment_K"]["seconds"]), "sweeps_and_K": n_pl,
                "cpu_sweep_s": sw["seconds"], "cpu_moment_K_s": c["moment_K"]["seconds"],
                "note": "lower bound: sweeps + moment-form K at the GPU solve leg's sweep count, so the ratio to "
                        "the GPU solve is not the dense -> moment change"}
    return out


def copy_calibration(gib=4):
    """A plain device copy of known size (for PMC byte-counter scaling) and its bandwidth."""
    import torch
    n = gib * (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del a, b
    torch.cuda.empty_cache()
    return {"bytes_read": gib * (1 << 30), "bytes_written": gib * (1 << 30),
            "copy_GBs": 2 * gib * (1 << 30) / dt / 1e9}


def settle_autotuner(op, m, ang, G=1):
    """Run the sweep autotuner to its choice on scratch vectors (part of setup,
    like the schedule build) and return the seconds it took."""
    import torch
    q = torch.rand(m.NE * ang.NI * 4 * G, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(q)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(32):
        if op.sweep_choice(G)[0] != "auto":
            break
        op.transport_matinv_dev(q, psi, 0, G)
    if G == 1:
        # the within-group solves' sweeps, which store the flux (and maybe read the source)
        # in sweep order, are timed apart
        for _ in range(32):
            if op.sweep_order_choice()[0] != "auto":
                break
            op.transport_matinv_order_dev(q, psi, 0)
        psi.copy_(q)
        for _ in range(32):
            if op.sweep_order_choice(True)[0] != "auto":
                break
            op.transport_matinv_order_dev(psi, psi, 0, source_in_order=True)
        # the FP32 flux mode's sweeps (the LDS kernel at its three workgroup sizes, two timings each)
        f32 = torch.empty(q.numel(), dtype=torch.float32, device="cuda")
        for _ in range(8):
            op.transport_matinv_order32_dev(q, psi, f32, 0)
        del f32
    op.synchronize()
    dt = time.perf_counter() - t0
    del q, psi
    torch.cuda.empty_cache()
    return dt


# Layouts of the within-group angular flux (Operator.set_flux_order); the library's default first.
FLUX_LAYOUTS = ("sweep", "sweep_source", "natural", "sweep_fp32")
DEFAULT_FLUX_LAYOUT = FLUX_LAYOUTS[0]


def solve_leg(m, ang, rtol, group_batch=1, layout=DEFAULT_FLUX_LAYOUT, others=True):
    """The 384,000-tet multigroup downscatter solve with the +z beam first-collision source;
    ``group_batch`` consecutive groups per coupled solve (1 = group by group), the within-group
    flux in ``layout`` (one group at a time only); ``others`` adds the same solve in the other
    layouts and with reduced-precision D psi."""
    import torch
    from lionbolt_amd import angular, problems
    from lionbolt_amd.ops import Operator
    G = 24
    sigt, S = problems.downscatter_xs(G=G, L=ang.L)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, problems.within_group(S))
    op.set_group_batch(group_batch)
    tune_s = settle_autotuner(op, m, ang, group_batch)
    if group_batch == 1:
        op.set_flux_order(layout)
    qmom, umom = _beam_qmom(op, m, sigt, S)
    phi = torch.zeros_like(qmom)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = op.multigroup_solve_dev(S, phi, qmom=qmom, kind="gmres", space="moment", m=30, rtol=rtol)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    choice = list(op.sweep_choice(group_batch))
    order_choice = {"flux": list(op.sweep_order_choice()), "source_and_flux": list(op.sweep_order_choice(True))}
    layouts = {}
    if group_batch == 1 and others:
        # the same solve with the other layouts of the within-group angular flux (bit-identical
        # moments between the FP64 ones)
        for order in FLUX_LAYOUTS:
            if order == layout:
                continue
            op.set_flux_order(order)
            phn = torch.zeros_like(qmom)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            sl = op.multigroup_solve_dev(S, phn, qmom=qmom, kind="gmres", space="moment", m=30, rtol=rtol)
            torch.cuda.synchronize()
            layouts[order] = {"seconds": time.perf_counter() - t1, "moments_bit_identical": bool(torch.equal(phn, phi))}
            if "sweep_fp32" in (order, layout):
                # the Arnoldi steps' flux for D in FP32: its error against the FP64 solve
                layouts[order].update({
                    "sweeps": sum(x["sweeps"] for x in sl), "iterates": [x["iterates"] for x in sl],
                    "max_true_residual": max(x["true_residual"] for x in sl),
                    "scalar_flux_max_rel_diff_vs_fp64":
                        float(((phn[:, 0] - phi[:, 0]).abs().amax(dim=1) / phi[:, 0].abs().amax(dim=1)).max())})
            del phn
        op.set_flux_order(layout)
    # scalar flux = sqrt(4 pi) x the l = 0 moment; group 1 adds its uncollided flux (Main.f90:443)
    s4pi = math.sqrt(4.0 * math.pi)
    flux = {"collided_sum_all_groups": float(phi[:, 0].sum()) * s4pi,
            "uncollided_sum_group1": float(umom[0].sum()) * s4pi}
    reduced = {}
    if group_batch == 1 and others and ang.NI == 256 and ang.L == 3:
        # D psi of the Arnoldi steps on MFMA with reduced inputs (FP32 accumulation);
        # right-hand side, psi and the true residual in FP64
        ref = phi[:, 0].clone()
        for fmt in ("split_bf16", "bf16"):
            op.set_moment_precision(fmt)
            phr = torch.zeros_like(qmom)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            sr = op.multigroup_solve_dev(S, phr, qmom=qmom, kind="gmres", space="moment", m=30, rtol=rtol,
                                         maxit=60, raise_on_fail=False)
            torch.cuda.synchronize()
            reduced[fmt] = {"seconds": time.perf_counter() - t1, "iterates": [x["iterates"] for x in sr],
                            "sweeps": sum(x["sweeps"] for x in sr),
                            "max_true_residual": max(x["true_residual"] for x in sr),
                            "groups_not_converged": sum(x["status"] != 0 for x in sr),
                            "scalar_flux_max_rel_diff_vs_fp64":
                                float(((phr[:, 0] - ref).abs().amax(dim=1) / ref.abs().amax(dim=1)).max())}
        op.set_moment_precision("fp64")
    op.close()
    max_res = max(s["true_residual"] for s in stats)
    return {"seconds": wall, "groups": G, "rtol": rtol, "autotune_s": tune_s, "sweep_choice": choice,
            "sweep_order_choice": order_choice,
            "flux_layout": layout if group_batch == 1 else "natural", "other_flux_layouts": layouts,
            "max_true_residual": max_res, "target_residual": 1e-6, "target_met": bool(max_res <= 1e-6),
            "iterates": [s["iterates"] for s in stats], "sweeps": sum(s["sweeps"] for s in stats),
            "group_batch": group_batch, "scalar_flux": flux, "reduced_precision_D_psi": reduced,
            "method": ("plain GMRES(30) on flux moments per group, FP64" if group_batch == 1 else
                       f"plain GMRES(30) on flux moments, {group_batch} groups per coupled system, FP64")}


def reference_method_leg(m, ang, rtol):
    """The reference's solver method on the GPU: psi-space GMRES(10) per group from
    x0 = q/sigma_t(1) (BoltzmannOperator.f90:106-151), same multigroup problem and
    stopping tolerance as the solve leg; K applied in moment form (the same
    operator as the dense kernel).  Gives the iteration counts the CPU projection
    uses, and the time of the reference method on the GPU."""
    import torch
    from lionbolt_amd import angular, problems
    from lionbolt_amd.ops import Operator
    G = 24
    sigt, S = problems.downscatter_xs(G=G, L=ang.L)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, problems.within_group(S))
    settle_autotuner(op, m, ang)
    qmom, _ = _beam_qmom(op, m, sigt, S)
    phi = torch.zeros_like(qmom)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = op.multigroup_solve_dev(S, phi, qmom=qmom, kind="gmres", space="psi", m=10, rtol=rtol)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    op.close()
    return {"seconds": wall, "rtol": rtol, "iterates": [s["iterates"] for s in stats],
            "sweeps": sum(s["sweeps"] for s in stats),
            "max_true_residual": max(s["true_residual"] for s in stats),
            "method": "psi-space GMRES(10) per group (the reference's method), FP64"}


def _beam_qmom(op, m, sigt, S):
    """Source moments of the +z beam per group (G x nmom x NENK), built as the
    reference builds its first-collision source (ray-traced uncollided fluence
    projected on the nodal basis, FCSSource.f90:134-215), and the moments of the
    beam group's uncollided angular fluence (nmom x NENK)."""
    from lionbolt_amd import problems
    qmom, _, umom = problems.beam_group_sources(op, m, S, sigt)
    return qmom, umom


def cube48k_leg():
    """The 48,000-tet problem end to end on the GPU through the host entry point (vectors cross PCIe):
    psi-space GMRES(30) at 1e-8 as the config states, and the moment-space solve."""
    from lionbolt_amd import problems
    from lionbolt_amd.ops import Operator
    m, ang, sigt, sigs, q = problems.cube_source_48k()
    op = Operator(m, ang)
    op.set_physics(sigt, sigs)
    op.boltzmann_matinv(q, kind="gmres", space="moment", m=30, rtol=1e-8)    # warm-up (autotuner, graphs)
    res = {}
    for space in ("psi", "moment"):
        t0 = time.perf_counter()
        _, st = op.boltzmann_matinv(q, kind="gmres", space=space, m=30, rtol=1e-8)
        res[space] = {"seconds": time.perf_counter() - t0, "iterates": st["iterates"],
                      "true_residual": st["true_residual"]}
    op.close()
    return res


def batch_leg(reps=10, rtol=1e-6):
    """64 uneven single-group S8 beam problems as one batch.  Sweep rate of the
    batched sweep per launch mode (the autotuned choice included), the same 64
    problems swept one operator at a time, and the batched per-problem GMRES."""
    import torch
    from lionbolt_amd import angular, problems
    from lionbolt_amd.ops import Operator
    um, eoff, sigt, sigs, probs = problems.uneven_beams_batch()
    ang = angular.level_symmetric_s8()
    ca = um.NE * ang.NI
    op = Operator(um, ang)
    op.set_physics(sigt, sigs)
    op.set_problems(eoff)
    qmom = problems.uneven_beams_sources(op, eoff, probs, sigt, sigs)
    q = torch.rand(op.n, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(q)

    def timed(k):
        op.transport_matinv_dev(q, psi, 0, 1)
        op.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            op.transport_matinv_dev(q, psi, 0, 1)
        op.synchronize()
        return 1000.0 * (time.perf_counter() - t0) / k

    res = {"problems": len(probs), "tets": um.NE, "directions": ang.NI,
           "tets_min": int(min(p["mesh"].NE for p in probs)), "tets_max": int(max(p["mesh"].NE for p in probs))}
    modes = {}
    op.set_sweep_mode("auto")
    for _ in range(32):
        if op.sweep_choice(1)[0] != "auto":
            break
        op.transport_matinv_dev(q, psi, 0, 1)
    op.synchronize()
    choice = op.sweep_choice(1)
    ms = timed(reps)
    modes["auto"] = {"ms_per_sweep": ms, "cell_angle_sweeps_per_s": ca / (ms * 1e-3), "choice": list(choice)}
    for name, mode, block, per_cu in (("levels", "levels", 0, 0), ("graph", "graph", 0, 0),
                                      ("persistent-1024x2", "persistent", 1024, 2),
                                      ("persistent-256x4", "persistent", 256, 4),
                                      ("persistent-256x8", "persistent", 256, 8),
                                      ("direction-1024", "direction", 1024, 0)):
        if block:
            op.set_direction_block(block)
        if per_cu:
            op.set_workgroups_per_cu(per_cu)
        op.set_sweep_mode(mode)
        ms = timed(reps)
        modes[name] = {"ms_per_sweep": ms, "cell_angle_sweeps_per_s": ca / (ms * 1e-3)}
    # the flux in the batch's sweep order, as the batched solve sweeps (autotuned persistent launch)
    op.set_sweep_mode("auto")
    for _ in range(32):
        if op.sweep_order_choice()[0] != "auto":
            break
        op.transport_matinv_order_dev(q, psi, 0)
    op.transport_matinv_order_dev(q, psi, 0)
    op.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        op.transport_matinv_order_dev(q, psi, 0)
    op.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    modes["auto-sweep-order"] = {"ms_per_sweep": ms, "cell_angle_sweeps_per_s": ca / (ms * 1e-3),
                                 "choice": list(op.sweep_order_choice())}
    res["sweep"] = modes
    # the same problems swept one operator at a time (each with its own autotuned launch)
    singles = []
    for pr in probs:
        o = Operator(pr["mesh"], ang)
        o.set_physics([pr["sigt"]], [[pr["c"] * pr["sigt"]]])
        qq = torch.rand(o.n, dtype=torch.float64, device="cuda")
        for _ in range(32):                 # through the autotuner's timing phase
            if o.sweep_choice(1)[0] != "auto":
                break
            o.transport_matinv_dev(qq, qq, 0, 1)
        o.transport_matinv_dev(qq, qq, 0, 1)
        singles.append((o, qq))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        for o, qq in singles:
            o.transport_matinv_dev(qq, qq, 0, 1)
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / 3
    res["one_operator_per_problem"] = {"ms_per_sweep_of_all": ms, "cell_angle_sweeps_per_s": ca / (ms * 1e-3)}
    for o, _ in singles:
        o.close()
    del singles
    # batched solve from the first-collision sources
    op.set_sweep_mode("auto")
    qm = qmom.reshape(-1).contiguous()
    phi = torch.zeros_like(qm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = op.batch_solve_dev(qm, phi, m=30, rtol=rtol)
    torch.cuda.synchronize()
    res["solve"] = {"seconds": time.perf_counter() - t0, "rtol": rtol,
                    "iterates_max": max(s["iterates"] for s in st),
                    "iterates_mean": float(np.mean([s["iterates"] for s in st])),
                    "sweeps": st[0]["sweeps"], "max_true_residual": max(s["true_residual"] for s in st),
                    "flux_layout": DEFAULT_FLUX_LAYOUT,
                    "method": "per-problem flux-moment GMRES(30) in lockstep, one sweep launch per step, FP64"}
    # the same solve with the angular flux in the natural layout (bit-identical moments)
    op.set_flux_order("natural")
    phn = torch.zeros_like(qm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.batch_solve_dev(qm, phn, m=30, rtol=rtol)
    torch.cuda.synchronize()
    res["solve"]["natural_flux_layout"] = {"seconds": time.perf_counter() - t0,
                                           "moments_bit_identical": bool(torch.equal(phn, phi))}
    op.close()
    return res


def learned_pc_leg(m, ang, rtol=1e-6, fmt="bf16"):
    """The 24-group P3 problem with within-group c = 0.99, solved group by
    group with plain flux-moment GMRES(30) and with flexible GMRES(30) right-
    preconditioned by the per-cell-node network, trained first on the solver's own
    residual/error pairs from 32 smaller seeded problems."""
    import torch
    from lionbolt_amd import learned_pc, problems
    from lionbolt_amd.ops import Operator
    X, Y, data = learned_pc.build_training_set()
    params, tinfo = learned_pc.train(X, Y)
    del X, Y
    t0 = time.perf_counter()
    thr, table = learned_pc.select_threshold(params, fmt=fmt)
    policy = {"threshold": thr if math.isfinite(thr) else "inf", "seconds": time.perf_counter() - t0,
              "training_problems_fgmres_faster": sum(r["fgmres_s"] < r["gmres_s"] for r in table),
              "training_problems": len(table),
              "thickness_range": [min(r["thickness"] for r in table), max(r["thickness"] for r in table)],
              "iterates_gmres_vs_fgmres": [[round(r["thickness"], 3), r["gmres_iterates"], r["fgmres_iterates"]]
                                           for r in sorted(table, key=lambda r: r["thickness"])]}
    G = 24
    sigt, S = problems.scattering_dominated_xs(G=G, L=ang.L)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, problems.within_group(S))
    op.set_preconditioner(params, fmt)
    op.set_preconditioner_threshold(thr)
    settle_autotuner(op, m, ang)
    qmom, _ = _beam_qmom(op, m, sigt, S)
    res = {"training": {"rows": data["rows"], "solves": len(data["solves"]),
                        "data_seconds": data["seconds"], **tinfo}, "policy": policy}
    for kind in ("gmres", "fgmres"):
        phi = torch.zeros_like(qmom)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = op.multigroup_solve_dev(S, phi, qmom=qmom, kind=kind, space="moment", m=30, rtol=rtol, atol=0.0)
        torch.cuda.synchronize()
        res[kind] = {"seconds": time.perf_counter() - t0, "iterates": [s["iterates"] for s in st],
                     "total_iterates": sum(s["iterates"] for s in st), "sweeps": sum(s["sweeps"] for s in st),
                     "max_true_residual": max(s["true_residual"] for s in st),
                     "group_seconds": [round(s["seconds"], 4) for s in st]}
        if kind == "fgmres":
            res[kind]["fallbacks"] = [s["fallback"] for s in st]   # -1: network not chosen for the group
            res[kind]["groups_with_network"] = sum(s["fallback"] != -1 for s in st)
            res[kind]["group_thickness"] = [round(float(sigt[g].max() * np.cbrt(m.volumes().sum())), 4)
                                            for g in range(G)]
            res[kind]["preconditioner"] = fmt
        if kind == "gmres":
            ref = phi.clone()
        else:
            res["flux_rel_diff"] = float((phi[:, 0] - ref[:, 0]).abs().max() / ref[:, 0].abs().max())
    res["speedup_fgmres_vs_gmres"] = res["gmres"]["seconds"] / res["fgmres"]["seconds"]
    # the network in every group (threshold 0, runtime fallback only): what the
    # a-priori choice avoids, group by group
    op.set_preconditioner_threshold(0.0)
    phi = torch.zeros_like(qmom)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = op.multigroup_solve_dev(S, phi, qmom=qmom, kind="fgmres", space="moment", m=30, rtol=rtol, atol=0.0)
    torch.cuda.synchronize()
    res["fgmres_network_in_every_group"] = {
        "seconds": time.perf_counter() - t0, "iterates": [s["iterates"] for s in st],
        "total_iterates": sum(s["iterates"] for s in st), "fallbacks": [s["fallback"] for s in st],
        "group_seconds": [round(s["seconds"], 4) for s in st],
        "gmres_group_seconds": res["gmres"].get("group_seconds"),
        "max_true_residual": max(s["true_residual"] for s in st)}
    op.set_preconditioner_threshold(thr)
    # one network apply on a full-size residual (RMS scaling + the MFMA kernel);
    # algorithmic bytes per spatial dof: 16 FP64 moments in and out, plus the
    # element's material and size shared by its 4 nodes (12 B / 4)
    r = torch.randn(op.nmom * op.NENK, dtype=torch.float64, device="cuda")
    z = torch.empty_like(r)
    op.mlp_apply_dev(r, z, 0)
    op.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        op.mlp_apply_dev(r, z, 0)
    op.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    res["network_apply"] = {"dofs": op.NENK, "ms": ms, "bytes_per_dof": 259,
                            "GBs_algorithmic": 259 * op.NENK / (ms * 1e-3) / 1e9, "format": fmt}
    op.close()
    return res


def high_order_leg(n=24, orders=(3, 7, 11, 15), reps=5):
    """The scattering orders of the reference's input decks (PNSCATTERING 7, 11,
    15; SNPrep sets of 128, 288, 512 directions) on an 82,944-tet cube, one
    group: one sweep and the two moment products of a flux-moment operator
    application, each timed alone after the autotuner settles."""
    import torch
    from lionbolt_amd import angular, mesh as M
    from lionbolt_amd.ops import Operator
    m = M.jittered_cube(n, a=30.0)
    out = {"tets": m.NE}
    for L in orders:
        ang = angular.snprep(L)
        op = Operator(m, ang)
        op.set_physics([0.5], [0.45 * 0.6 ** np.arange(L + 1)])
        q = torch.rand(m.NE * 4 * ang.NI, dtype=torch.float64, device="cuda")
        psi = torch.empty_like(q)
        phi = torch.empty(ang.n_moments * m.NENK, dtype=torch.float64, device="cuda")
        settle_autotuner(op, m, ang)

        def timed(f):
            f()
            op.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                f()
            op.synchronize()
            return 1000.0 * (time.perf_counter() - t0) / reps

        ts = timed(lambda: op.transport_matinv_dev(q, psi, 0, 1))
        td = timed(lambda: op.moments_dev(psi, phi))
        tm = timed(lambda: op.moments_to_discrete_dev(phi, q))
        flops = 2.0 * ang.n_moments * ang.NI * m.NENK
        out[f"L{L}"] = {"directions": ang.NI, "moments": ang.n_moments, "sweep_ms": ts, "d_psi_ms": td,
                        "m_phi_ms": tm, "d_psi_tflops": flops / td / 1e9, "m_phi_tflops": flops / tm / 1e9,
                        "sweep_choice": list(op.sweep_choice(1))}
        op.close()
        del q, psi, phi
        torch.cuda.empty_cache()
    return out


def single_group_leg(m, ang, reps=10, only=None):
    """One-group sweeps of the same mesh and directions (what a within-group solve
    issues), per launch mode: milliseconds per sweep and cell-angle sweeps/s.
    ``only``: the names of the variants to run (all when None), so that a counter
    run sees one variant of a kernel name."""
    import torch
    from lionbolt_amd import problems
    from lionbolt_amd.ops import Operator
    sigt, _ = problems.downscatter_xs(G=1, L=ang.L)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, np.zeros((1, 2, ang.L + 1)))
    q = torch.rand(m.NE * ang.NI * 4, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(q)
    res = {}
    for mode in ("levels", "graph", "direction", "direction-256", "direction-512", "direction_lds-256",
                 "direction_lds-512", "auto"):
        if only is not None and mode not in only:
            continue
        if mode == "auto":
            op.set_sweep_mode("auto")
            for _ in range(32):
                if op.sweep_choice(1)[0] != "auto":
                    break
                op.transport_matinv_dev(q, psi, 0, 1)
        elif mode.startswith("direction"):
            op.set_direction_block(int(mode.split("-")[1]) if "-" in mode else 1024)
            op.set_sweep_mode(mode.split("-")[0])
        else:
            op.set_sweep_mode(mode)
        op.transport_matinv_dev(q, psi, 0, 1)
        op.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            op.transport_matinv_dev(q, psi, 0, 1)
        op.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0) / reps
        res[mode] = {"ms_per_sweep": ms, "cell_angle_sweeps_per_s": m.NE * ang.NI / (ms * 1e-3)}
        if mode == "auto":
            res[mode]["choice"] = list(op.sweep_choice(1))
    # in place, as the natural-layout solve sweeps; and with the flux stored in sweep
    # order (out of place), as the flux-moment solve sweeps
    op.set_direction_block(768)
    op.set_sweep_mode("direction_lds")
    for name, run in (("direction_lds-768-in-place", lambda: op.transport_matinv_dev(psi, psi, 0, 1)),
                      ("direction_lds-768-sweep-order", lambda: op.transport_matinv_order_dev(q, psi, 0)),
                      ("direction_lds-768-sweep-order-in-place",
                       lambda: op.transport_matinv_order_dev(psi, psi, 0, source_in_order=True)),
                      ("direction_lds-512-sweep-order", None), ("direction-512-sweep-order", None),
                      ("levels-sweep-order", None)):
        if only is not None and name not in only:
            continue
        if run is None:
            mode = name.rsplit("-sweep-order", 1)[0]
            if "-" in mode:
                op.set_direction_block(int(mode.split("-")[1]))
            op.set_sweep_mode(mode.split("-")[0])
            run = lambda: op.transport_matinv_order_dev(q, psi, 0)   # noqa: E731
        run()
        op.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        op.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0) / reps
        res[name] = {"ms_per_sweep": ms, "cell_angle_sweeps_per_s": m.NE * ang.NI / (ms * 1e-3)}
    op.close()
    return res


def _note(msg):
    """A progress line on stderr (the JSON result alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run(args):
    import torch
    from lionbolt_amd import problems
    from lionbolt_amd.ops import Operator

    torch.cuda.set_device(0)
    out_extra = {}
    if args.calib:
        out_extra["copy_calibration"] = copy_calibration()
    t0 = time.perf_counter()
    m = problems.slab_cube_mesh(args.n)
    ang = problems.product_8x32()
    G = args.groups
    sigt, _ = problems.downscatter_xs(G=G, L=ang.L)
    op = Operator(m, ang)
    op.set_physics_groups(sigt, np.zeros((G, 2, ang.L + 1)))
    op.set_sweep_mode(args.sweep_mode)
    setup_s = time.perf_counter() - t0
    cag = m.NE * ang.NI * G
    q = torch.rand(cag * 4, dtype=torch.float64, device="cuda")
    psi = torch.empty_like(q)
    torch.cuda.synchronize()
    mode, block = args.sweep_mode, 0
    if args.sweep_mode == "auto":
        # the autotuner times each launch configuration on its first sweeps (untimed here)
        for _ in range(32):
            if op.sweep_choice(G)[0] != "auto":
                break
            op.transport_matinv_dev(q, psi, 0, G)
        op.synchronize()
        mode, block, _ = op.sweep_choice(G)
    elif args.sweep_mode in ("direction", "direction_lds"):
        block = args.block
        op.set_direction_block(block)
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
    ms_step = 1000.0 * wall / max(args.steps, 1)
    value = cag * args.steps / wall
    ev_ms = float(np.mean(sweep_ms)) if sweep_ms else float("nan")
    achieved = BYTES_PER_CAG * cag / (ev_ms * 1e-3) / 1e9
    if not bool(torch.isfinite(psi[:: 1 << 16]).all()):
        raise RuntimeError("non-finite flux")
    other_modes = {}
    if args.compare_modes:
        for other in ("levels", "graph", "direction", "direction_lds"):
            if other == args.sweep_mode:
                continue
            op.set_sweep_mode(other)
            op.transport_matinv_dev(q, psi, 0, G)
            op.synchronize()
            t1 = time.perf_counter()
            for _ in range(3):
                op.transport_matinv_dev(q, psi, 0, G)
            op.synchronize()
            other_modes[other] = {"ms_per_sweep": 1000.0 * (time.perf_counter() - t1) / 3}
    levels = op.maxlev
    launches = 1 if mode.startswith("direction") else levels
    kernel = {"direction": f"sweep_direction_kernel<{block}>",
              "direction_lds": f"sweep_direction_lds_kernel<{block}>"}.get(mode, "sweep_level_kernel")
    op.close()
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
                               f"P{ang.L}, in HBM",
                   "tets": m.NE, "directions": ang.NI, "groups": G, "levels_max": levels,
                   "sweep_mode": args.sweep_mode, "launch": {"mode": mode, "block": block},
                   "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": kernel, "launches_per_sweep": launches, "bytes_per_unit": BYTES_PER_CAG,
                     "event_ms_per_sweep": ev_ms},
    }
    traffic_file = os.path.join(ROOT, "profiles", "sweep_traffic.json")
    tr = None
    if os.path.exists(traffic_file) and args.n == 40 and G == 24:
        with open(traffic_file) as fh:
            tr = json.load(fh).get(kernel)
    if tr:
        # HBM bytes per launch from the two PMC passes (FETCH_SIZE, WRITE_SIZE) recorded there
        per_cag = tr["fetch_bytes_per_cag"] + tr["write_bytes_per_cag"]
        out["roofline"]["traffic"] = per_cag * cag / launches
        out["roofline"]["traffic_unit"] = "bytes per launch (PMC, " + tr["source"] + ")"
        out["roofline"]["traffic_bytes_per_unit"] = per_cag
    out.update(out_extra)
    _note(f"headline sweep: {ms_step:.2f} ms per step, {value:.4g} cell-angle-groups/s")
    if other_modes:
        out["other_sweep_modes"] = other_modes
    if args.single:
        out["single_group_sweep"] = single_group_leg(m, ang, only=args.single_modes)
        _note("single-group sweep modes done")
    if args.high_order:
        out["high_order_operator"] = high_order_leg()
        _note("scattering orders 3-15 done")
    if args.solve:
        runs = [solve_leg(m, ang, args.solve_rtol, b, args.flux_layout, args.solve_variants)
                for b in args.group_batches]
        _note("24-group solve done")
        out["solve"] = min(runs, key=lambda r: r["seconds"])
        out["solve_by_group_batch"] = {str(r["group_batch"]): {k: r[k] for k in ("seconds", "sweeps",
                                                                                "max_true_residual")}
                                       for r in runs}
    ref = None
    if args.ref_method:
        out["reference_method_gpu"] = reference_method_leg(m, ang, 1e-6)
        _note("reference method done")
        ref = {"sweeps": out["reference_method_gpu"]["sweeps"],
               "plain_sweeps": out["solve"]["sweeps"] if "solve" in out else None}
    if args.cube48k:
        out["cube48k_gpu"] = cube48k_leg()
        _note("48,000-tet solve done")
    if args.batch:
        out["uneven_beams_batch"] = batch_leg()
        _note("64-problem batch done")
    if args.learned_pc:
        out["learned_pc"] = learned_pc_leg(m, ang)
        _note("learned preconditioner done")
    if args.gemm_study:
        from lionbolt_amd import gemm_study
        out["moment_gemm_study"] = gemm_study.run(ncols=m.NENK, groups=G)
    if args.cpu_baseline:
        _note("CPU baseline (oracle on the host cores) ...")
        out["cpu_baseline"] = cpu_baseline(args.n, ref)
        _note("CPU baseline done")
        for key in ("solve_projected", "solve_projected_moment_form"):
            proj = out["cpu_baseline"].get(key)
            if proj and "solve" in out:
                proj["gpu_solve_s"] = out["solve"]["seconds"]
                proj["speedup_vs_gpu_solve"] = proj["seconds"] / out["solve"]["seconds"]
        if "cube48k_gpu" in out and "cube48k_solve" in out["cpu_baseline"]:
            out["cube48k_gpu"]["speedup_vs_cpu"] = out["cpu_baseline"]["cube48k_solve"]["seconds"] / out["cube48k_gpu"]["psi"]["seconds"]
    print(json.dumps(out))


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=40, help="hexes per side (40 -> 384,000 tets)")
    p.add_argument("--groups", type=int, default=24)
    p.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    p.add_argument("--no-solve", dest="solve", action="store_false")
    p.add_argument("--flux-layout", choices=FLUX_LAYOUTS, default=DEFAULT_FLUX_LAYOUT,
                   help="layout of the within-group flux in the timed solve (default: the library's)")
    p.add_argument("--no-solve-variants", dest="solve_variants", action="store_false",
                   help="time only the solve in --flux-layout (no other layouts, no reduced-precision D psi)")
    p.add_argument("--no-single", dest="single", action="store_false")
    p.add_argument("--single-modes", type=lambda t: t.split(","), default=None,
                   help="comma-separated one-group sweep variants to run (default: all)")
    p.add_argument("--no-high-order", dest="high_order", action="store_false",
                   help="skip the scattering orders 7, 11 and 15 of the reference's decks")
    p.add_argument("--solve-rtol", type=float, default=1e-6,
                   help="GMRES tolerance of the solve leg; its true transport residual must reach 1e-6")
    p.add_argument("--group-batches", type=lambda t: [int(x) for x in t.split(",")], default=[1],
                   help="energy groups per coupled solve to try in the solve leg (e.g. 1,2,4)")
    p.add_argument("--no-ref-method", dest="ref_method", action="store_false",
                   help="skip the reference-method (psi-space GMRES(10)) solve on the GPU")
    p.add_argument("--no-cube48k", dest="cube48k", action="store_false")
    p.add_argument("--no-batch", dest="batch", action="store_false", help="skip the 64-problem batch")
    p.add_argument("--no-learned-pc", dest="learned_pc", action="store_false",
                   help="skip the learned-preconditioner solve")
    p.add_argument("--calib", action="store_true", help="run a known-size copy first (PMC scaling)")
    p.add_argument("--sweep-mode", default="auto",
                   choices=["auto", "levels", "graph", "direction", "direction_lds"])
    p.add_argument("--block", type=int, default=1024, choices=[256, 512, 768, 1024],
                   help="workgroup size of the direction-resident sweep modes")
    p.add_argument("--compare-modes", action="store_true")
    p.add_argument("--gemm-study", action="store_true", help="scattering-moment products on MFMA, all formats")
    args = p.parse_args()
    if int(os.environ.get("WORLD_SIZE", "1")) > 1 or args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (no collective path is built)"}))
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
