"""Scattering-moment products on MFMA against FP64 on the same operands.

The discrete-to-moment D psi and moment-to-discrete M phi run in FP32, BF16,
split BF16 (hi + lo) and scaled FP16 with FP32 accumulation.  Error bounds per
output element, relative to its column's largest magnitude, follow each format's
unit roundoff (FP32 2^-24, BF16 2^-8, split BF16 about 2^-16, FP16 2^-11) times
a modest growth factor; a wrong operand layout would give O(1) errors."""
import numpy as np
import pytest

from lionbolt_amd import angular, gemm_study

pytestmark = pytest.mark.gpu

BOUNDS = {"fp32": 2e-6, "bf16": 4e-2, "split_bf16": 2e-4, "fp16_scaled": 5e-3}


@pytest.mark.parametrize("fmt", list(gemm_study.FORMATS))
def test_moment_products_match_fp64(fmt):
    import torch
    ang = angular.product(8, 32, 3)
    D64, M64 = angular.moment_matrices(ang)
    Dd = torch.from_numpy(np.ascontiguousarray(D64)).cuda()
    Md = torch.from_numpy(np.ascontiguousarray(M64)).cuda()
    ncols = 4096 + 7
    psi = gemm_study.lognormal_psi(ncols, seed=1003)
    ref_phi = Dd @ psi
    phi = gemm_study.gemm_d(fmt, Dd, gemm_study.pack(fmt, psi))
    mx, rms = gemm_study.rel_errors(phi, ref_phi, 0)
    assert mx <= BOUNDS[fmt], (fmt, mx)
    assert rms <= BOUNDS[fmt] / 4
    ref_psi = (Md @ ref_phi).T
    out = gemm_study.gemm_m(fmt, Md, gemm_study.pack(fmt, ref_phi))
    mx2, rms2 = gemm_study.rel_errors(out, ref_psi, 1)
    assert mx2 <= BOUNDS[fmt], (fmt, mx2)


def test_precision_ordering():
    import torch
    ang = angular.product(8, 32, 3)
    D64, _ = angular.moment_matrices(ang)
    Dd = torch.from_numpy(np.ascontiguousarray(D64)).cuda()
    psi = gemm_study.lognormal_psi(2048, seed=7)
    ref = Dd @ psi
    err = {f: gemm_study.rel_errors(gemm_study.gemm_d(f, Dd, gemm_study.pack(f, psi)), ref, 0)[1]
           for f in gemm_study.FORMATS}
    assert err["fp32"] < err["split_bf16"] < err["fp16_scaled"] < err["bf16"]
