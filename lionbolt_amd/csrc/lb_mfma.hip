// Scattering-moment products on MFMA (gfx950), in reduced input precision with
// FP32 accumulation:
//   discrete -> moment  phi[m][col] = sum_i D[m][i] psi[i][col]   (K = NI)
//   moment -> discrete  psi[col][i] = sum_m M[i][m] phi[m][col]   (K = nmom)
// Input formats (LB_FMT_*): FP32; BF16; split BF16 (x = hi + lo, both BF16,
// products hi*hi + hi*lo + lo*hi); FP16 with a power-of-two scale per column.
// Reduced-precision operands are stored with the reduced dimension contiguous per
// column ("column-major"): psi as [col][NI], phi as [col][nmom], so each lane's
// MFMA operand is one contiguous run.  The matrices D and M arrive in FP64 and are
// rounded to the input format inside the kernel.
//
// The discrete->moment product is HBM-bound (256 x 2 B in + 16 x 4 B out per
// column in BF16: 8192 flops per 576 B); the MFMA rate needed is a few percent of
// the chip's.  Lane layout of v_mfma_f32_16x16x{4,16,32}: the lanes l and l' with
// l % 16 equal share a column of B and a row of A; lane group kb = l / 16 supplies
// the kb-th block of the K slots of both operands, and the 16x16 FP32 result holds
// rows 4 kb .. 4 kb + 3 of column l % 16 in lane l.  Because A and B share the K
// slots of each lane, K can be assigned to lanes in any order: each lane group
// takes one contiguous quarter of the reduced dimension.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "lb_internal.hpp"

namespace lb {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(4 * sizeof(short)))) short s16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

enum { FMT_FP32 = 0, FMT_BF16 = 1, FMT_SPLIT_BF16 = 2, FMT_FP16 = 3 };

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

__device__ __forceinline__ short bf16_bits(__bf16 x) { return __builtin_bit_cast(short, x); }

// ---------------------------------------------------------------------------
// Packing: src [rows][ncols] FP64 -> dst [ncols][rows] in the format, through an
// LDS tile so both the read and the write are coalesced.  For FP16 the caller's
// per-column exponent e (a power of two 2^e brings the column's largest magnitude
// to [2^13, 2^14)) is applied; inv_scale[col] = 2^-e.
// ---------------------------------------------------------------------------
template <int FMT>
__global__ __launch_bounds__(256) void pack_kernel(int rows, int64_t ncols, const double* __restrict__ src,
                                                   void* __restrict__ dst, void* __restrict__ dst_lo,
                                                   const float* __restrict__ inv_scale) {
  __shared__ double tile[32][33];
  const int64_t c0 = (int64_t)blockIdx.x * 32;
  const int r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k;
    const int64_t c = c0 + tx;
    tile[k][tx] = (r < rows && c < ncols) ? src[(int64_t)r * ncols + c] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t c = c0 + k;
    const int r = r0 + tx;
    if (c >= ncols || r >= rows) continue;
    const double v = tile[tx][k];
    const int64_t o = c * rows + r;
    if (FMT == FMT_FP32) {
      static_cast<float*>(dst)[o] = (float)v;
    } else if (FMT == FMT_BF16) {
      static_cast<__bf16*>(dst)[o] = to_bf16((float)v);
    } else if (FMT == FMT_SPLIT_BF16) {
      const __bf16 hi = to_bf16((float)v);
      static_cast<__bf16*>(dst)[o] = hi;
      static_cast<__bf16*>(dst_lo)[o] = to_bf16((float)(v - (double)(float)hi));
    } else {
      static_cast<_Float16*>(dst)[o] = (_Float16)(v / (double)inv_scale[c]);
    }
  }
}

// ---------------------------------------------------------------------------
// discrete -> moment: one wave per 16-column tile (grid-stride), NI % 64 == 0,
// nmom <= 16.  Lane group kb owns directions [kb NI/4, (kb+1) NI/4).
// ---------------------------------------------------------------------------
template <int FMT, int NI>
__global__ __launch_bounds__(256) void d_gemm_kernel(int nmom, const double* __restrict__ D, int64_t ncols,
                                                     const void* __restrict__ psi, const void* __restrict__ psi_lo,
                                                     const float* __restrict__ inv_scale, float* __restrict__ phi) {
  const int lane = threadIdx.x & 63;
  const int j = lane & 15, kb = lane >> 4;
  constexpr int Q = NI / 4;                 // directions per lane group
  const int row = j;                        // A row (moment) held by this lane
  const bool row_ok = row < nmom;
  const int64_t waves = (int64_t)gridDim.x * 4;
  const int64_t ntile = (ncols + 15) / 16;
  // A operand of the format, converted once per wave
  constexpr int NA = FMT == FMT_FP32 ? Q : Q / 8;
  typedef typename std::conditional<FMT == FMT_FP32, float,
                                    typename std::conditional<FMT == FMT_FP16, f16x8, bf16x8>::type>::type AT;
  AT a[NA], alo[FMT == FMT_SPLIT_BF16 ? NA : 1];
  if constexpr (FMT == FMT_FP32) {
#pragma unroll
    for (int s = 0; s < Q; ++s) a[s] = row_ok ? (float)D[row * NI + kb * Q + s] : 0.f;
  } else {
#pragma unroll
    for (int s = 0; s < Q / 8; ++s)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const double d = row_ok ? D[row * NI + kb * Q + 8 * s + e] : 0.0;
        if constexpr (FMT == FMT_FP16) {
          a[s][e] = (_Float16)d;
        } else {
          const __bf16 hi = to_bf16((float)d);
          a[s][e] = hi;
          if constexpr (FMT == FMT_SPLIT_BF16) alo[s][e] = to_bf16((float)(d - (double)(float)hi));
        }
      }
  }
  // B operand of one tile: this lane's Q directions of column t * 16 + j (and
  // the low halves for split BF16), loaded one tile ahead of its products so
  // that a wave keeps two tiles' loads in flight.
  constexpr int NB = FMT == FMT_FP32 ? Q / 4 : Q / 8;
  typedef typename std::conditional<FMT == FMT_FP32, float4, AT>::type BT;
  struct Tile {
    BT v[NB];
    BT vl[FMT == FMT_SPLIT_BF16 ? NB : 1];
  };
  auto load = [&](int64_t t, Tile& T) {
    const int64_t col = t * 16 + j;
    const bool ok = col < ncols;
    const int64_t o = col * NI + kb * Q;
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      if constexpr (FMT == FMT_FP32) {
        T.v[s] = ok ? reinterpret_cast<const float4*>(static_cast<const float*>(psi) + o)[s]
                    : make_float4(0.f, 0.f, 0.f, 0.f);
      } else if constexpr (FMT == FMT_FP16) {
        T.v[s] = ok ? reinterpret_cast<const f16x8*>(static_cast<const _Float16*>(psi) + o)[s] : f16x8{};
      } else {
        T.v[s] = ok ? reinterpret_cast<const bf16x8*>(static_cast<const __bf16*>(psi) + o)[s] : bf16x8{};
        if constexpr (FMT == FMT_SPLIT_BF16)
          T.vl[s] = ok ? reinterpret_cast<const bf16x8*>(static_cast<const __bf16*>(psi_lo) + o)[s] : bf16x8{};
      }
    }
  };
  auto product = [&](const Tile& T) {
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NB; ++s) {
      if constexpr (FMT == FMT_FP32) {
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * s + 0], T.v[s].x, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * s + 1], T.v[s].y, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * s + 2], T.v[s].z, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * s + 3], T.v[s].w, c, 0, 0, 0);
      } else if constexpr (FMT == FMT_FP16) {
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[s], T.v[s], c, 0, 0, 0);
      } else if constexpr (FMT == FMT_BF16) {
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], T.v[s], c, 0, 0, 0);
      } else {
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[s], T.v[s], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], T.vl[s], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], T.v[s], c, 0, 0, 0);
      }
    }
    return c;
  };
  int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  Tile cur;
  if (t < ntile) load(t, cur);
  while (t < ntile) {
    const int64_t tn = t + waves;
    Tile nxt;
    if (tn < ntile) load(tn, nxt);
    const f32x4 c = product(cur);
    const int64_t col = t * 16 + j;
    if (col < ncols) {
      const float sc = FMT == FMT_FP16 ? inv_scale[col] : 1.f;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (4 * kb + r < nmom) phi[(int64_t)(4 * kb + r) * ncols + col] = FMT == FMT_FP16 ? c[r] * sc : c[r];
    }
    cur = nxt;
    t = tn;
  }
}

// ---------------------------------------------------------------------------
// moment -> discrete: nmom = 16, NI % 16 == 0.  B = phi tile (16 moments x 16
// columns, lane group kb holds moments 4 kb .. 4 kb + 3 of column j); A = M rows of
// one 16-direction tile; result lanes write 4 consecutive directions of a column.
// ---------------------------------------------------------------------------
template <int FMT, int NI>
__global__ __launch_bounds__(256) void m_gemm_kernel(const double* __restrict__ M, int64_t ncols,
                                                     const void* __restrict__ phi, const void* __restrict__ phi_lo,
                                                     const float* __restrict__ inv_scale, float* __restrict__ psi) {
  const int lane = threadIdx.x & 63;
  const int j = lane & 15, kb = lane >> 4;
  constexpr int NT = NI / 16;
  const int64_t waves = (int64_t)gridDim.x * 4;
  const int64_t ntile = (ncols + 15) / 16;
  if (FMT == FMT_FP32) {
    float a[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) a[t][s] = (float)M[(16 * t + j) * 16 + 4 * kb + s];
    for (int64_t tt = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); tt < ntile; tt += waves) {
      const int64_t col = tt * 16 + j;
      const bool ok = col < ncols;
      const float4 v = ok ? *reinterpret_cast<const float4*>(static_cast<const float*>(phi) + col * 16 + 4 * kb)
                          : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][0], v.x, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][1], v.y, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][2], v.z, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][3], v.w, c, 0, 0, 0);
        if (ok) *reinterpret_cast<f32x4*>(psi + col * NI + 16 * t + 4 * kb) = c;
      }
    }
  } else {
    // 16x16x16 BF16 / 16x16x32 FP16 both fit nmom = 16 with one K block of 4 per lane
    // group (FP16 uses the x32 form with the upper K half zero).
    s16x4 a[NT], alo[NT];
    f16x8 ah[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double d = M[(16 * t + j) * 16 + 4 * kb + s];
        const __bf16 hi = to_bf16((float)d);
        a[t][s] = bf16_bits(hi);
        alo[t][s] = bf16_bits(to_bf16((float)(d - (double)(float)hi)));
        ah[t][s] = (_Float16)d;
        ah[t][s + 4] = (_Float16)0.f;
      }
    for (int64_t tt = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); tt < ntile; tt += waves) {
      const int64_t col = tt * 16 + j;
      const bool ok = col < ncols;
      const int64_t o = col * 16 + 4 * kb;
      s16x4 v = {}, vl = {};
      f16x8 vh = {};
      float sc = 1.f;
      if (ok) {
        if (FMT == FMT_FP16) {
          const _Float16* p = static_cast<const _Float16*>(phi) + o;
#pragma unroll
          for (int s = 0; s < 4; ++s) vh[s] = p[s];
          sc = inv_scale[col];
        } else {
          v = *reinterpret_cast<const s16x4*>(static_cast<const __bf16*>(phi) + o);
          if (FMT == FMT_SPLIT_BF16) vl = *reinterpret_cast<const s16x4*>(static_cast<const __bf16*>(phi_lo) + o);
        }
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        if (FMT == FMT_FP16) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[t], vh, c, 0, 0, 0);
          c[0] *= sc;
          c[1] *= sc;
          c[2] *= sc;
          c[3] *= sc;
        } else if (FMT == FMT_BF16) {
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], v, c, 0, 0, 0);
        } else {
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo[t], v, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], vl, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], v, c, 0, 0, 0);
        }
        if (ok) *reinterpret_cast<f32x4*>(psi + col * NI + 16 * t + 4 * kb) = c;
      }
    }
  }
}

// per-column FP16 exponent: 2^-e with e chosen so max |x| 2^e lies in [2^13, 2^14)
__global__ __launch_bounds__(256) void fp16_scale_kernel(int rows, int64_t ncols, const double* __restrict__ src,
                                                         float* __restrict__ inv_scale) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= ncols) return;
  double mx = 0.0;
  for (int r = 0; r < rows; ++r) mx = fmax(mx, fabs(src[(int64_t)r * ncols + c]));
  int e = 0;
  if (mx > 0.0) {
    int ex;
    frexp(mx, &ex);          // mx = f 2^ex, f in [0.5, 1)
    e = 14 - ex;             // mx 2^e in [2^13, 2^14)
  }
  inv_scale[c] = ldexpf(1.0f, -e);
}

// Grid of the two products: one wave per 16-column tile up to `cap`
// workgroups, grid-stride beyond.  Each wave converts its operand of D or M
// once.  The M product (one 16-direction tile of M per wave, 16 B loads of phi)
// runs best with 2048 workgroups (8 per CU), each wave walking many tiles
// (BF16 at 6.1 M columns: 3.0 -> 4.0 TB/s); the D product, whose tile loop
// waits on its own loads, needs the extra waves of up to 65,536 workgroups to
// keep bytes in flight (2,048 workgroups: 2.6 -> 1.4 TB/s at 36.9 M columns).
int grid_cols(int64_t ncols, int64_t cap) {
  const int64_t tiles = (ncols + 15) / 16;
  const int64_t blocks = (tiles + 3) / 4;
  return (int)(blocks < cap ? (blocks > 0 ? blocks : 1) : cap);
}

__global__ __launch_bounds__(256) void widen_kernel(int64_t n, const float* __restrict__ in, double* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) out[j] = in[j];
}

}  // namespace

}  // namespace lb

using namespace lb;

extern "C" {
int lb_pack_dev(int fmt, int rows, int64_t ncols, const double* src, void* dst, void* dst_lo, float* inv_scale,
                void* stream);
int lb_gemm_d_dev(int fmt, int nmom, int NI, const double* D, int64_t ncols, const void* psi, const void* psi_lo,
                  const float* inv_scale, float* phi, void* stream);
}

namespace lb {

// D psi inside the flux-moment GMRES with reduced-precision inputs: psi
// [NI][NENK] is packed column-major in the format (lb_pack_dev), multiplied on
// MFMA with FP32 accumulation (lb_gemm_d_dev) and the FP32 moments are widened
// into phi [nmom][NENK].
int launch_moments_reduced(lb_op* op, const double* psi, double* phi) {
  const int fmt = op->moment_fmt, NI = (int)op->NI, nmom = (int)op->nmom;
  const int64_t ncols = op->NENK;
  const size_t elem = fmt == FMT_FP32 ? 4 : 2;
  const size_t hi = (size_t)NI * ncols * elem, lo = fmt == FMT_SPLIT_BF16 ? (size_t)NI * ncols * 2 : 0;
  const size_t sc = fmt == FMT_FP16 ? (size_t)ncols * 4 : 0, out = (size_t)nmom * ncols * 4;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  LB_TRY(op->w_mix.reserve(al(hi) + al(lo) + al(sc) + al(out)));
  char* base = op->w_mix.as<char>();
  void* d_hi = base;
  void* d_lo = lo ? base + al(hi) : nullptr;
  float* d_sc = sc ? reinterpret_cast<float*>(base + al(hi) + al(lo)) : nullptr;
  float* d_out = reinterpret_cast<float*>(base + al(hi) + al(lo) + al(sc));
  LB_TRY(lb_pack_dev(fmt, NI, ncols, psi, d_hi, d_lo, d_sc, op->stream));
  LB_TRY(lb_gemm_d_dev(fmt, nmom, NI, op->d_D.as<const double>(), ncols, d_hi, d_lo, d_sc, d_out, op->stream));
  const int64_t n = (int64_t)nmom * ncols;
  hipLaunchKernelGGL(widen_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)), dim3(256), 0,
                     op->stream, n, d_out, phi);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

}  // namespace lb

extern "C" {

int lb_pack_dev(int fmt, int rows, int64_t ncols, const double* src, void* dst, void* dst_lo, float* inv_scale,
                void* stream) {
  return guarded([&]() -> int {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (rows < 1 || ncols < 1 || !src || !dst || (fmt == FMT_SPLIT_BF16 && !dst_lo) || (fmt == FMT_FP16 && !inv_scale) ||
        fmt < 0 || fmt > 3) {
      set_error("lb_pack_dev: invalid argument");
      return LB_E_ARG;
    }
    if (fmt == FMT_FP16)
      hipLaunchKernelGGL(fp16_scale_kernel, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0, s, rows, ncols, src,
                         inv_scale);
    dim3 grid((unsigned)((ncols + 31) / 32), (unsigned)((rows + 31) / 32));
    switch (fmt) {
      case FMT_FP32: hipLaunchKernelGGL(pack_kernel<FMT_FP32>, grid, dim3(256), 0, s, rows, ncols, src, dst, dst_lo, inv_scale); break;
      case FMT_BF16: hipLaunchKernelGGL(pack_kernel<FMT_BF16>, grid, dim3(256), 0, s, rows, ncols, src, dst, dst_lo, inv_scale); break;
      case FMT_SPLIT_BF16: hipLaunchKernelGGL(pack_kernel<FMT_SPLIT_BF16>, grid, dim3(256), 0, s, rows, ncols, src, dst, dst_lo, inv_scale); break;
      default: hipLaunchKernelGGL(pack_kernel<FMT_FP16>, grid, dim3(256), 0, s, rows, ncols, src, dst, dst_lo, inv_scale); break;
    }
    LB_HIP(hipGetLastError());
    return LB_OK;
  });
}

int lb_gemm_d_dev(int fmt, int nmom, int NI, const double* D, int64_t ncols, const void* psi, const void* psi_lo,
                  const float* inv_scale, float* phi, void* stream) {
  return guarded([&]() -> int {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (NI != 256 || nmom < 1 || nmom > 16 || ncols < 1 || !D || !psi || !phi || fmt < 0 || fmt > 3 ||
        (fmt == FMT_SPLIT_BF16 && !psi_lo) || (fmt == FMT_FP16 && !inv_scale)) {
      set_error("lb_gemm_d_dev: needs NI = 256, nmom <= 16 and the operands of the format");
      return LB_E_ARG;
    }
    const int g = grid_cols(ncols, 8192);
    switch (fmt) {
      case FMT_FP32: hipLaunchKernelGGL((d_gemm_kernel<FMT_FP32, 256>), dim3(g), dim3(256), 0, s, nmom, D, ncols, psi, psi_lo, inv_scale, phi); break;
      case FMT_BF16: hipLaunchKernelGGL((d_gemm_kernel<FMT_BF16, 256>), dim3(g), dim3(256), 0, s, nmom, D, ncols, psi, psi_lo, inv_scale, phi); break;
      case FMT_SPLIT_BF16: hipLaunchKernelGGL((d_gemm_kernel<FMT_SPLIT_BF16, 256>), dim3(g), dim3(256), 0, s, nmom, D, ncols, psi, psi_lo, inv_scale, phi); break;
      default: hipLaunchKernelGGL((d_gemm_kernel<FMT_FP16, 256>), dim3(g), dim3(256), 0, s, nmom, D, ncols, psi, psi_lo, inv_scale, phi); break;
    }
    LB_HIP(hipGetLastError());
    return LB_OK;
  });
}

int lb_gemm_m_dev(int fmt, int NI, int nmom, const double* M, int64_t ncols, const void* phi, const void* phi_lo,
                  const float* inv_scale, float* psi, void* stream) {
  return guarded([&]() -> int {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (NI != 256 || nmom != 16 || ncols < 1 || !M || !phi || !psi || fmt < 0 || fmt > 3 ||
        (fmt == FMT_SPLIT_BF16 && !phi_lo) || (fmt == FMT_FP16 && !inv_scale)) {
      set_error("lb_gemm_m_dev: needs NI = 256, nmom = 16 and the operands of the format");
      return LB_E_ARG;
    }
    const int g = grid_cols(ncols, (fmt == FMT_BF16 || fmt == FMT_SPLIT_BF16) ? 4096 : 65536);
    switch (fmt) {
      case FMT_FP32: hipLaunchKernelGGL((m_gemm_kernel<FMT_FP32, 256>), dim3(g), dim3(256), 0, s, M, ncols, phi, phi_lo, inv_scale, psi); break;
      case FMT_BF16: hipLaunchKernelGGL((m_gemm_kernel<FMT_BF16, 256>), dim3(g), dim3(256), 0, s, M, ncols, phi, phi_lo, inv_scale, psi); break;
      case FMT_SPLIT_BF16: hipLaunchKernelGGL((m_gemm_kernel<FMT_SPLIT_BF16, 256>), dim3(g), dim3(256), 0, s, M, ncols, phi, phi_lo, inv_scale, psi); break;
      default: hipLaunchKernelGGL((m_gemm_kernel<FMT_FP16, 256>), dim3(g), dim3(256), 0, s, M, ncols, phi, phi_lo, inv_scale, psi); break;
    }
    LB_HIP(hipGetLastError());
    return LB_OK;
  });
}

}  // extern "C"
