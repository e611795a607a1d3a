"""Scattering-moment products on MFMA in reduced input precision.

Discrete-to-moment phi = D psi (D: 16 x 256 real spherical harmonics times
weights of the GL8 x C32 set) and moment-to-discrete psi = M phi (M: 256 x 16),
per group over 1,536,000 cell-nodes, with psi drawn log-normal (mu = 0,
sigma = 2).  Formats: FP32, BF16, split BF16 (hi + lo, three products) and FP16
with a power-of-two scale per column, all accumulating in FP32.  Errors are
measured against the FP64 product on the same operands, per output element
relative to the largest magnitude of its column.
"""
import math
import time

import numpy as np

from . import _lib, angular

FORMATS = {"fp32": 0, "bf16": 1, "split_bf16": 2, "fp16_scaled": 3}
IN_BYTES = {"fp32": 4, "bf16": 2, "split_bf16": 4, "fp16_scaled": 2}
PEAK = {"fp32": 157.3e12, "bf16": 2.5e15, "split_bf16": 2.5e15, "fp16_scaled": 2.5e15}


def _ptr(t):
    return None if t is None else t.data_ptr()


def pack(fmt, src):
    """src (rows, ncols) FP64 cuda tensor -> dict of the format's operands ((ncols, rows) storage)."""
    import torch
    rows, ncols = src.shape
    dt = {"fp32": torch.float32, "bf16": torch.bfloat16, "split_bf16": torch.bfloat16,
          "fp16_scaled": torch.float16}[fmt]
    main = torch.empty((ncols, rows), dtype=dt, device=src.device)
    lo = torch.empty_like(main) if fmt == "split_bf16" else None
    sc = torch.empty(ncols, dtype=torch.float32, device=src.device) if fmt == "fp16_scaled" else None
    _lib.check(_lib.lib().lb_pack_dev(FORMATS[fmt], rows, ncols, src.data_ptr(), main.data_ptr(), _ptr(lo),
                                      _ptr(sc), None))
    return {"main": main, "lo": lo, "scale": sc, "ncols": ncols}


def gemm_d(fmt, D, packed, out=None):
    import torch
    ncols = packed["ncols"]
    nmom = D.shape[0]
    if out is None:
        out = torch.empty((nmom, ncols), dtype=torch.float32, device=D.device)
    _lib.check(_lib.lib().lb_gemm_d_dev(FORMATS[fmt], nmom, D.shape[1], D.data_ptr(), ncols, packed["main"].data_ptr(),
                                        _ptr(packed["lo"]), _ptr(packed["scale"]), out.data_ptr(), None))
    return out


def gemm_m(fmt, M, packed, out=None):
    import torch
    ncols = packed["ncols"]
    if out is None:
        out = torch.empty((ncols, M.shape[0]), dtype=torch.float32, device=M.device)
    _lib.check(_lib.lib().lb_gemm_m_dev(FORMATS[fmt], M.shape[0], M.shape[1], M.data_ptr(), ncols,
                                        packed["main"].data_ptr(), _ptr(packed["lo"]), _ptr(packed["scale"]),
                                        out.data_ptr(), None))
    return out


def rel_errors(got, ref, col_axis):
    """max and RMS of |got - ref| / max_col |ref| (per column of the output)."""
    import torch
    ref = ref.to(torch.float64)
    scale = ref.abs().amax(dim=col_axis, keepdim=True).clamp_min(1e-300)
    e = (got.to(torch.float64) - ref).abs() / scale
    return float(e.max()), float(torch.sqrt((e * e).mean()))


def lognormal_psi(ncols, seed, ni=256):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.exp(2.0 * torch.randn((ni, ncols), dtype=torch.float64, device="cuda", generator=g))


def run(ncols=1536000, groups=24, seed=1003, reps=3, formats=tuple(FORMATS)):
    """Errors on one group, timings over all groups: a dict per format and direction."""
    import torch
    ang = angular.product(8, 32, 3)
    D64, M64 = angular.moment_matrices(ang)
    Dd = torch.from_numpy(np.ascontiguousarray(D64)).cuda()
    Md = torch.from_numpy(np.ascontiguousarray(M64)).cuda()
    psi = lognormal_psi(ncols, seed)
    ref_phi = Dd @ psi                                    # FP64 reference, (16, ncols)
    ref_psi = (Md @ ref_phi).T.contiguous()               # FP64 reference of M phi, (ncols, 256)
    out = {}
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ntot = ncols * groups
    for fmt in formats:
        res = {}
        pk = pack(fmt, psi)
        phi = gemm_d(fmt, Dd, pk)
        res["d_max_rel_err"], res["d_rms_rel_err"] = rel_errors(phi, ref_phi, 0)
        pkm = pack(fmt, ref_phi)
        pso = gemm_m(fmt, Md, pkm)
        res["m_max_rel_err"], res["m_rms_rel_err"] = rel_errors(pso, ref_psi, 1)
        del pk, phi, pkm, pso
        torch.cuda.empty_cache()
        # timing at full size: all groups in one launch, operands packed from
        # regenerated log-normal groups (the values do not change the time)
        big = {"ncols": ntot}
        dt = {"fp32": torch.float32, "fp16_scaled": torch.float16}.get(fmt, torch.bfloat16)
        big["main"] = torch.empty((ntot, 256), dtype=dt, device="cuda")
        big["lo"] = torch.empty_like(big["main"]) if fmt == "split_bf16" else None
        big["scale"] = torch.ones(ntot, dtype=torch.float32, device="cuda") if fmt == "fp16_scaled" else None
        for gi in range(groups):
            sl = slice(gi * ncols, (gi + 1) * ncols)
            part = pack(fmt, psi)
            big["main"][sl] = part["main"]
            if big["lo"] is not None:
                big["lo"][sl] = part["lo"]
            if big["scale"] is not None:
                big["scale"][sl] = part["scale"]
            del part
        phi_all = torch.empty((16, ntot), dtype=torch.float32, device="cuda")
        gemm_d(fmt, Dd, big, phi_all)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            gemm_d(fmt, Dd, big, phi_all)
        ev1.record()
        torch.cuda.synchronize()
        t_d = ev0.elapsed_time(ev1) / 1000.0 / reps
        del big
        torch.cuda.empty_cache()
        # moment -> discrete at full size
        bigm = {"ncols": ntot}
        bigm["main"] = torch.zeros((ntot, 16), dtype=dt, device="cuda")
        bigm["lo"] = torch.zeros_like(bigm["main"]) if fmt == "split_bf16" else None
        bigm["scale"] = torch.ones(ntot, dtype=torch.float32, device="cuda") if fmt == "fp16_scaled" else None
        psi_all = torch.empty((ntot, 256), dtype=torch.float32, device="cuda")
        gemm_m(fmt, Md, bigm, psi_all)
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(reps):
            gemm_m(fmt, Md, bigm, psi_all)
        ev1.record()
        torch.cuda.synchronize()
        t_m = ev0.elapsed_time(ev1) / 1000.0 / reps
        del bigm, psi_all, phi_all
        torch.cuda.empty_cache()
        b = IN_BYTES[fmt]
        d_bytes = (256 * b + 16 * 4 + (4 if fmt == "fp16_scaled" else 0)) * ntot
        m_bytes = (16 * b + 256 * 4 + (4 if fmt == "fp16_scaled" else 0)) * ntot
        flops = 2 * 16 * 256 * ntot * (3 if fmt == "split_bf16" else 1)
        res.update({"d_seconds": t_d, "d_GBs": d_bytes / t_d / 1e9, "d_hbm_frac": d_bytes / t_d / 8e12,
                    "d_mfma_util": flops / t_d / PEAK[fmt],
                    "m_seconds": t_m, "m_GBs": m_bytes / t_m / 1e9, "m_hbm_frac": m_bytes / t_m / 8e12,
                    "m_mfma_util": flops / t_m / PEAK[fmt]})
        out[fmt] = res
    # FP64 baseline of the same products (library GEMM in FP64, for comparison)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        _ = Dd @ psi
    torch.cuda.synchronize()
    out["fp64_torch_per_group_d_seconds"] = (time.perf_counter() - t0) / reps
    return out


def main():
    """Command-line run of the study (the profile summaries under profiles/ use it)."""
    import argparse
    import json
    p = argparse.ArgumentParser(description="scattering-moment products on MFMA: error and speed per format")
    p.add_argument("--ncols", type=int, default=1536000, help="cell-nodes per group")
    p.add_argument("--groups", type=int, default=24)
    p.add_argument("--reps", type=int, default=3)
    p.add_argument("--formats", default=",".join(FORMATS))
    p.add_argument("--out", default=None, help="also write the result JSON here")
    a = p.parse_args()
    res = run(ncols=a.ncols, groups=a.groups, reps=a.reps, formats=tuple(a.formats.split(",")))
    text = json.dumps(res)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text)
    print(text)


if __name__ == "__main__":
    main()
