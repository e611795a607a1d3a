// Learned right preconditioner: a per-cell-node MLP 18 -> 64 -> 64 -> 64 -> 16
// (tanh on the hidden layers) that maps a residual's 16 flux moments at one
// spatial dof, log(sigma_t h) and sigma_s0 / sigma_t of the dof's element, to an
// estimate of the error's 16 moments there.  The residual is normalised by its
// root-mean-square s before the network and the output is scaled back by s, so
// P(a r) = a P(r) for a > 0.  No reference counterpart: the reference has plain
// GMRES and source iteration only.
//
// Parameters (FP32, 10,576): W1 [18][64], b1 [64], W2 [64][64], b2, W3 [64][64], b3,
// W4 [64][16], b4 [16]; W[i][j] connects input i to output j.
//
// Inference runs on MFMA (v_mfma_f32_16x16x32_bf16): a wave takes 16 dofs, the
// weights are the A operand (rows = output neurons), the activations the B
// operand (columns = dofs).  The result of a layer leaves lane l holding neurons
// 16 t + 4 (l / 16) + r (tiles t = 0..3, r = 0..3) of dof l % 16; since A and B
// only need to agree on which input neuron each K slot carries, those registers
// are fed to the next layer as they are and the weights are packed in the same
// slot order (host side, once).  An FP32 path (one thread per dof, weights in
// LDS) serves as the precision reference.
//
// Training (FP32): minibatch Adam over (feature, target) rows; the forward and
// backward GEMMs run on FP32 matrix cores (v_mfma_f32_16x16x4_f32), the weight
// gradients' sums over the minibatch split 16 ways and reduced in a fixed order,
// so gradients are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "lb_internal.hpp"

namespace lb {

namespace {

constexpr int kIn = 18, kH = 64, kOut = 16;
constexpr int kW1 = 0, kB1 = kW1 + kIn * kH, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kH * kH, kW4 = kB3 + kH, kB4 = kW4 + kH * kOut, kParams = kB4 + kOut;
static_assert(kParams == LB_MLP_PARAMS, "parameter count");
// packed BF16 weights, [tile][step][lane][8]
constexpr int kP1 = 0, kP2 = kP1 + 4 * 1 * 64 * 8, kP3 = kP2 + 4 * 2 * 64 * 8, kP4 = kP3 + 4 * 2 * 64 * 8,
              kPackLen = kP4 + 1 * 2 * 64 * 8;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

struct FeatArgs {
  const double* r;         // [16][NENK] residual moments (single group)
  int64_t NENK;
  double inv_s;            // 1 / rms(r)
  const int32_t* mat;      // [NE]
  const double* sigt;      // [Nm][Gtot]
  const double* sigs;      // [Nm][Gtot][L+1]
  int Gtot, g0, L;
  const double* elem_h;    // [NE]
  // when set, {1 / rms(r), output scale} computed on the device replace inv_s and
  // the kernels' scale argument (the allocation-free entry point)
  const double* dev_scale = nullptr;
};

__device__ __forceinline__ double input_scale(const FeatArgs& F) { return F.dev_scale ? F.dev_scale[0] : F.inv_s; }
__device__ __forceinline__ double output_scale(const FeatArgs& F, double s) { return F.dev_scale ? F.dev_scale[1] : s; }

__device__ __forceinline__ float feature(const FeatArgs& F, int64_t sd, int f, double inv_s) {
  if (f < 16) return (float)(F.r[(int64_t)f * F.NENK + sd] * inv_s);
  const int64_t e = sd >> 2;
  const int m = F.mat[e];
  const double st = F.sigt[(int64_t)m * F.Gtot + F.g0];
  if (f == 16) return (float)log(st * F.elem_h[e]);
  if (f == 17) return (float)(F.sigs[((int64_t)m * F.Gtot + F.g0) * (F.L + 1)] / st);
  return 0.f;
}

__global__ __launch_bounds__(256) void features_kernel(FeatArgs F, float* __restrict__ out) {
  const int64_t sd = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (sd >= F.NENK) return;
  const double inv_s = input_scale(F);
#pragma unroll
  for (int f = 0; f < kIn; ++f) out[sd * kIn + f] = feature(F, sd, f, inv_s);
}

__device__ __forceinline__ float act(float x) { return tanhf(x); }

// tanh x = 1 - 2 / (e^{2x} + 1) from the hardware's base-2 exponential and
// reciprocal (e^{2x} = 2^{2x / ln 2}): within a few 1e-7 of tanhf in absolute
// terms (near 0 the subtraction cancels, so the error is absolute, not
// relative), far below the BF16 rounding (2^-9 of the largest activations) of
// the values it feeds.  Saturates to +-1 for large |x| (the exponential
// overflows to inf or underflows to 0).  The BF16 kernel only: training and the
// FP32 path keep tanhf.
__device__ __forceinline__ float act_bf16(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
  return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

// FP32 reference inference: one thread per dof.
__global__ __launch_bounds__(256) void mlp_fp32_kernel(FeatArgs F, const float* __restrict__ W, double s,
                                                       double* __restrict__ z) {
  __shared__ float w[kParams];
  for (int i = threadIdx.x; i < kParams; i += 256) w[i] = W[i];
  __syncthreads();
  const int64_t sd = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (sd >= F.NENK) return;
  const double inv_s = input_scale(F);
  s = output_scale(F, s);
  float x[kIn], h[kH], g[kH];
#pragma unroll
  for (int f = 0; f < kIn; ++f) x[f] = feature(F, sd, f, inv_s);
  for (int j = 0; j < kH; ++j) {
    float a = w[kB1 + j];
#pragma unroll
    for (int i = 0; i < kIn; ++i) a += x[i] * w[kW1 + i * kH + j];
    h[j] = act(a);
  }
  for (int j = 0; j < kH; ++j) {
    float a = w[kB2 + j];
    for (int i = 0; i < kH; ++i) a += h[i] * w[kW2 + i * kH + j];
    g[j] = act(a);
  }
  for (int j = 0; j < kH; ++j) {
    float a = w[kB3 + j];
    for (int i = 0; i < kH; ++i) a += g[i] * w[kW3 + i * kH + j];
    h[j] = act(a);
  }
  for (int j = 0; j < kOut; ++j) {
    float a = w[kB4 + j];
    for (int i = 0; i < kH; ++i) a += h[i] * w[kW4 + i * kOut + j];
    z[(int64_t)j * F.NENK + sd] = s * (double)a;
  }
}

// BF16 MFMA inference: each wave loops over tiles of 16 dofs.
__global__ __launch_bounds__(256) void mlp_bf16_kernel(FeatArgs F, const __bf16* __restrict__ P,
                                                       const float* __restrict__ W, double s,
                                                       double* __restrict__ z) {
  const int lane = threadIdx.x & 63;
  const int kb = lane >> 4, col = lane & 15;
  const int wave = threadIdx.x >> 6;
  const bf16x8* pk = reinterpret_cast<const bf16x8*>(P);
  bf16x8 w1[4], w2[4][2], w3[4][2], w4[2];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    w1[t] = pk[(kP1 >> 3) + t * 64 + lane];
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      w2[t][st] = pk[(kP2 >> 3) + (t * 2 + st) * 64 + lane];
      w3[t][st] = pk[(kP3 >> 3) + (t * 2 + st) * 64 + lane];
    }
  }
#pragma unroll
  for (int st = 0; st < 2; ++st) w4[st] = pk[(kP4 >> 3) + st * 64 + lane];
  float b1[4][4], b2[4][4], b3[4][4], b4[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      b1[t][r] = W[kB1 + 16 * t + 4 * kb + r];
      b2[t][r] = W[kB2 + 16 * t + 4 * kb + r];
      b3[t][r] = W[kB3 + 16 * t + 4 * kb + r];
    }
#pragma unroll
  for (int r = 0; r < 4; ++r) b4[r] = W[kB4 + 4 * kb + r];

  const double inv_s = input_scale(F);
  s = output_scale(F, s);
  const int64_t ntiles = (F.NENK + 15) / 16;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wave; tile < ntiles; tile += (int64_t)gridDim.x * 4) {
    const int64_t sd = tile * 16 + col;
    const bool ok = sd < F.NENK;
    bf16x8 xb;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int f = 8 * kb + q;
      xb[q] = (__bf16)((ok && f < kIn) ? feature(F, sd, f, inv_s) : 0.f);
    }
    f32x4 a[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      a[t] = f32x4{b1[t][0], b1[t][1], b1[t][2], b1[t][3]};
      a[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1[t], xb, a[t], 0, 0, 0);
    }
    bf16x8 hb[2];
#pragma unroll
    for (int st = 0; st < 2; ++st)
#pragma unroll
      for (int q = 0; q < 8; ++q) hb[st][q] = (__bf16)act_bf16(a[2 * st + (q >> 2)][q & 3]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      a[t] = f32x4{b2[t][0], b2[t][1], b2[t][2], b2[t][3]};
      a[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2[t][0], hb[0], a[t], 0, 0, 0);
      a[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2[t][1], hb[1], a[t], 0, 0, 0);
    }
#pragma unroll
    for (int st = 0; st < 2; ++st)
#pragma unroll
      for (int q = 0; q < 8; ++q) hb[st][q] = (__bf16)act_bf16(a[2 * st + (q >> 2)][q & 3]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      a[t] = f32x4{b3[t][0], b3[t][1], b3[t][2], b3[t][3]};
      a[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3[t][0], hb[0], a[t], 0, 0, 0);
      a[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3[t][1], hb[1], a[t], 0, 0, 0);
    }
#pragma unroll
    for (int st = 0; st < 2; ++st)
#pragma unroll
      for (int q = 0; q < 8; ++q) hb[st][q] = (__bf16)act_bf16(a[2 * st + (q >> 2)][q & 3]);
    f32x4 o = f32x4{b4[0], b4[1], b4[2], b4[3]};
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w4[0], hb[0], o, 0, 0, 0);
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w4[1], hb[1], o, 0, 0, 0);
    if (ok) {
#pragma unroll
      for (int r = 0; r < 4; ++r) z[(int64_t)(4 * kb + r) * F.NENK + sd] = s * (double)o[r];
    }
  }
}

// ---- training GEMMs on FP32 matrix cores -----------------------------------
//
// C (M x N) = sum_k A(m, k) B(k, n) with v_mfma_f32_16x16x4_f32: one wave per
// 16 x 16 tile of C, k in steps of 4 (lane l supplies A(m0 + l % 16, k0 + l / 16)
// and B(k0 + l / 16, n0 + l % 16) and receives C(m0 + 4 (l / 16) + t, n0 + l % 16),
// t < 4).  A(m, k) = a[ga(m) sam + gk(k) sak] with optional row gathers ga / gk
// (the minibatch rows of the training set), B(k, n) = b[k sbk + n sbn].  With
// splits > 1 the k range is cut into `splits` pieces whose partial tiles land in
// `part` and are summed in split order by gemm_reduce_kernel, so the sums do not
// depend on scheduling.  Epilogues: kBias (+ bias[n], tanh if act), kTanhGrad
// (x (1 - h(m, n)^2), the tanh derivative of the layer below), kPlain.
enum { kBias = 0, kTanhGrad = 1, kPlain = 2 };

struct GemmArgs {
  int M, N, K;
  const float* a;
  const int32_t* ga;   // gather of A's rows (m), or null
  const int32_t* gk;   // gather of A's columns (k) as rows of the stored array, or null
  int64_t sam, sak;
  const float* b;
  int64_t sbk, sbn;
  float* c;
  int64_t ldc;
  int mode, act;
  const float* bias;
  const float* h;      // kTanhGrad: [M][ldc]
  int splits;
  float* part;         // [splits][M][N] when splits > 1
};

typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4v;

__device__ __forceinline__ float gemm_epilogue(const GemmArgs& g, int i, int j, float v) {
  if (g.mode == kBias) {
    if (g.bias) v += g.bias[j];
    if (g.act) v = tanhf(v);
  } else if (g.mode == kTanhGrad) {
    const float hv = g.h[(int64_t)i * g.ldc + j];
    v *= 1.f - hv * hv;
  }
  return v;
}

__global__ __launch_bounds__(256) void mfma_gemm_kernel(GemmArgs g) {
  const int lane = threadIdx.x & 63;
  const int tn = (g.N + 15) / 16, tm = (g.M + 15) / 16;
  const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wid >= tm * tn * g.splits) return;   // whole waves leave together
  const int split = wid / (tm * tn), tile = wid - split * tm * tn;
  const int m0 = (tile / tn) * 16, n0 = (tile % tn) * 16;
  const int r = lane & 15, q = lane >> 4;
  const int kchunk = ((g.K + g.splits - 1) / g.splits + 3) / 4 * 4;
  const int k_begin = split * kchunk, k_end = min(g.K, k_begin + kchunk);
  const int am = m0 + r, bn = n0 + r;
  const bool a_ok = am < g.M, b_ok = bn < g.N;
  const int64_t arow = a_ok ? (int64_t)(g.ga ? g.ga[am] : am) * g.sam : 0;
  f32x4v acc = {0.f, 0.f, 0.f, 0.f};
  int k0 = k_begin;
  // eight k-steps' operands are loaded before their products
  for (; k0 + 32 <= k_end; k0 += 32) {
    float av[8], bv[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int k = k0 + 4 * s + q;
      av[s] = a_ok ? g.a[arow + (int64_t)(g.gk ? g.gk[k] : k) * g.sak] : 0.f;
      bv[s] = b_ok ? g.b[(int64_t)k * g.sbk + (int64_t)bn * g.sbn] : 0.f;
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], bv[s], acc, 0, 0, 0);
  }
  for (; k0 < k_end; k0 += 4) {
    const int k = k0 + q;
    const bool k_ok = k < k_end;
    const float av = (a_ok && k_ok) ? g.a[arow + (int64_t)(g.gk ? g.gk[k] : k) * g.sak] : 0.f;
    const float bv = (b_ok && k_ok) ? g.b[(int64_t)k * g.sbk + (int64_t)bn * g.sbn] : 0.f;
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
  }
  const int j = n0 + r;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = m0 + 4 * q + t;
    if (i < g.M && j < g.N) {
      if (g.splits > 1)
        g.part[((int64_t)split * g.M + i) * g.N + j] = acc[t];
      else
        g.c[(int64_t)i * g.ldc + j] = gemm_epilogue(g, i, j, acc[t]);
    }
  }
}

__global__ __launch_bounds__(256) void gemm_reduce_kernel(GemmArgs g) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)g.M * g.N) return;
  const int i = int(t / g.N), j = int(t - (int64_t)i * g.N);
  float v = 0.f;
  for (int s = 0; s < g.splits; ++s) v += g.part[((int64_t)s * g.M + i) * g.N + j];
  g.c[(int64_t)i * g.ldc + j] = gemm_epilogue(g, i, j, v);
}

// db[j] = sum_b dout[b][j] (J divides 1024): thread t sums column t % J over rows
// t / J, t / J + S, ... (S = 1024 / J slices, rows in order), then the slices are
// added pairwise in a fixed tree, so the sum's bits do not vary from run to run.
constexpr int kBiasThreads = 1024;
__global__ __launch_bounds__(kBiasThreads) void bias_grad_kernel(int B, int J, const float* __restrict__ dout,
                                                                 float* __restrict__ db) {
  __shared__ float sh[kBiasThreads];
  const int t = threadIdx.x, S = kBiasThreads / J;
  const int j = t % J, r = t / J;
  float a = 0.f;
  for (int b = r; b < B; b += S) a += dout[(int64_t)b * J + j];
  sh[t] = a;
  __syncthreads();
  for (int h = S / 2; h > 0; h >>= 1) {
    if (r < h) sh[t] += sh[t + h * J];
    __syncthreads();
  }
  if (r == 0) db[j] = sh[t];
}
static_assert(kBiasThreads % kH == 0 && kBiasThreads % kOut == 0 && (kH & (kH - 1)) == 0 && (kOut & (kOut - 1)) == 0,
              "bias-gradient slices must be a power of two");

// ---- training kernels (FP32) ------------------------------------------------

// dy = 2 (yhat - y) / (B * kOut); part[block] = sum of squared errors / (B * kOut)
__global__ __launch_bounds__(256) void loss_grad_kernel(int B, const float* __restrict__ yhat,
                                                        const float* __restrict__ Y, const int32_t* __restrict__ idx,
                                                        float* __restrict__ dy, float* __restrict__ part) {
  __shared__ float sh[256];
  const int t = blockIdx.x * 256 + threadIdx.x;
  float e2 = 0.f;
  if (t < B * kOut) {
    const int b = t / kOut, j = t - b * kOut;
    const float d = yhat[t] - Y[(int64_t)idx[b] * kOut + j];
    dy[t] = 2.f * d / (float)(B * kOut);
    e2 = d * d / (float)(B * kOut);
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void sum_kernel(int n, const float* __restrict__ part, float* __restrict__ out) {
  if (threadIdx.x == 0) {
    float a = 0.f;
    for (int i = 0; i < n; ++i) a += part[i];
    out[0] = a;
  }
}

__global__ __launch_bounds__(256) void adam_kernel(int n, float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                   float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

uint16_t bf16_rne(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Partial tiles of a split-k training GEMM: splits x M x N floats at most.
constexpr int kSplits = 16;
constexpr size_t kPartFloats = (size_t)kSplits * kH * kH;

void launch_gemm(hipStream_t st, const GemmArgs& g) {
  const int tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
  hipLaunchKernelGGL(mfma_gemm_kernel, dim3((unsigned)((tiles * g.splits + 3) / 4)), dim3(256), 0, st, g);
  if (g.splits > 1)
    hipLaunchKernelGGL(gemm_reduce_kernel, dim3((unsigned)(((int64_t)g.M * g.N + 255) / 256)), dim3(256), 0, st, g);
}

GemmArgs gemm_args(int M, int N, int K, const float* a, const int32_t* ga, const int32_t* gk, int64_t sam,
                   int64_t sak, const float* b, int64_t sbk, int64_t sbn, float* c, int64_t ldc, int mode, int act,
                   const float* bias, const float* h, int splits, float* part) {
  GemmArgs g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.a = a;
  g.ga = ga;
  g.gk = gk;
  g.sam = sam;
  g.sak = sak;
  g.b = b;
  g.sbk = sbk;
  g.sbn = sbn;
  g.c = c;
  g.ldc = ldc;
  g.mode = mode;
  g.act = act;
  g.bias = bias;
  g.h = h;
  g.splits = splits;
  g.part = part;
  return g;
}

// Activations, output errors and gradients of one minibatch step.
struct StepBufs {
  float *h1, *h2, *h3, *yh, *d3, *d2, *d1, *dy, *part, *lpart, *gpart;
};

// Forward pass, loss and its gradient with respect to every parameter for the
// rows idx[0..B-1] of (X, Y): the network's GEMMs on FP32 matrix cores, the
// weight gradients (sums over the minibatch, k = B) split 16 ways.
void enqueue_step(hipStream_t st, int B, const float* X, const float* Y, const int32_t* idx, const float* w,
                  float* g, float* loss, const StepBufs& t) {
  const int sp = B >= 256 ? kSplits : 1;
  // forward: h = tanh(in W + b), yh = h3 W4 + b4
  launch_gemm(st, gemm_args(B, kH, kIn, X, idx, nullptr, kIn, 1, w + kW1, kH, 1, t.h1, kH, kBias, 1, w + kB1,
                            nullptr, 1, nullptr));
  launch_gemm(st, gemm_args(B, kH, kH, t.h1, nullptr, nullptr, kH, 1, w + kW2, kH, 1, t.h2, kH, kBias, 1, w + kB2,
                            nullptr, 1, nullptr));
  launch_gemm(st, gemm_args(B, kH, kH, t.h2, nullptr, nullptr, kH, 1, w + kW3, kH, 1, t.h3, kH, kBias, 1, w + kB3,
                            nullptr, 1, nullptr));
  launch_gemm(st, gemm_args(B, kOut, kH, t.h3, nullptr, nullptr, kH, 1, w + kW4, kOut, 1, t.yh, kOut, kBias, 0,
                            w + kB4, nullptr, 1, nullptr));
  const unsigned lb = (unsigned)(((int64_t)B * kOut + 255) / 256);
  hipLaunchKernelGGL(loss_grad_kernel, dim3(lb), dim3(256), 0, st, B, t.yh, Y, idx, t.dy, t.lpart);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, st, (int)lb, t.lpart, loss);
  // backward: dW = in^T d (k = minibatch rows), db = column sums of d, d_in = (d W^T) (1 - h^2)
  launch_gemm(st, gemm_args(kH, kOut, B, t.h3, nullptr, nullptr, 1, kH, t.dy, kOut, 1, g + kW4, kOut, kPlain, 0,
                            nullptr, nullptr, sp, t.part));
  hipLaunchKernelGGL(bias_grad_kernel, dim3(1), dim3(kBiasThreads), 0, st, B, kOut, t.dy, g + kB4);
  launch_gemm(st, gemm_args(B, kH, kOut, t.dy, nullptr, nullptr, kOut, 1, w + kW4, 1, kOut, t.d3, kH, kTanhGrad, 0,
                            nullptr, t.h3, 1, nullptr));
  launch_gemm(st, gemm_args(kH, kH, B, t.h2, nullptr, nullptr, 1, kH, t.d3, kH, 1, g + kW3, kH, kPlain, 0, nullptr,
                            nullptr, sp, t.part));
  hipLaunchKernelGGL(bias_grad_kernel, dim3(1), dim3(kBiasThreads), 0, st, B, kH, t.d3, g + kB3);
  launch_gemm(st, gemm_args(B, kH, kH, t.d3, nullptr, nullptr, kH, 1, w + kW3, 1, kH, t.d2, kH, kTanhGrad, 0, nullptr,
                            t.h2, 1, nullptr));
  launch_gemm(st, gemm_args(kH, kH, B, t.h1, nullptr, nullptr, 1, kH, t.d2, kH, 1, g + kW2, kH, kPlain, 0, nullptr,
                            nullptr, sp, t.part));
  hipLaunchKernelGGL(bias_grad_kernel, dim3(1), dim3(kBiasThreads), 0, st, B, kH, t.d2, g + kB2);
  launch_gemm(st, gemm_args(B, kH, kH, t.d2, nullptr, nullptr, kH, 1, w + kW2, 1, kH, t.d1, kH, kTanhGrad, 0, nullptr,
                            t.h1, 1, nullptr));
  launch_gemm(st, gemm_args(kIn, kH, B, X, nullptr, idx, 1, kIn, t.d1, kH, 1, g + kW1, kH, kPlain, 0, nullptr,
                            nullptr, sp, t.part));
  hipLaunchKernelGGL(bias_grad_kernel, dim3(1), dim3(kBiasThreads), 0, st, B, kH, t.d1, g + kB1);
}

// Carves the step buffers out of `f` (floats): (3 kH + kOut) B activations,
// (3 kH + kOut) B errors, the split partials and 256 loss partials.
StepBufs carve(float* f, int B) {
  StepBufs t;
  t.h1 = f;
  t.h2 = t.h1 + (size_t)B * kH;
  t.h3 = t.h2 + (size_t)B * kH;
  t.yh = t.h3 + (size_t)B * kH;
  t.d3 = t.yh + (size_t)B * kOut;
  t.d2 = t.d3 + (size_t)B * kH;
  t.d1 = t.d2 + (size_t)B * kH;
  t.dy = t.d1 + (size_t)B * kH;
  t.part = t.dy + (size_t)B * kOut;
  t.lpart = t.part + kPartFloats;
  t.gpart = t.lpart + 256;
  return t;
}
constexpr size_t step_floats(int B) { return (size_t)B * (3 * kH + kOut) * 2 + kPartFloats + 256; }

// Input neuron carried by K slot q of lane group kb in step st: layer 1 reads
// the features in order; deeper layers read the registers the layer below left.
int slot_input(int layer, int kb, int st, int q) {
  if (layer == 1) return 8 * kb + q;
  return 16 * (2 * st + (q >> 2)) + 4 * kb + (q & 3);
}

// Element sizes h = (6 V)^(1/3), uploaded once per operator.
int ensure_elem_h(lb_op* op) {
  if (op->d_elem_h.p) return LB_OK;
  std::vector<double> h(op->NE);
  for (int64_t e = 0; e < op->NE; ++e) h[e] = std::cbrt(op->h_detJ[e]);
  return upload(op, op->d_elem_h, h.data(), h.size() * sizeof(double));
}

// Network inputs of residual r for group g0, except the residual's scale.
FeatArgs feat_base(lb_op* op, const double* r, int g0) {
  return FeatArgs{r, op->NENK, 0.0, op->d_mat.as<const int32_t>(), op->d_sigt.as<const double>(),
                  op->d_sigs.as<const double>(), op->G, g0, op->L, op->d_elem_h.as<const double>()};
}

int feat_args(lb_op* op, const double* r, int g0, FeatArgs* F) {
  if (op->nmom != 16) {
    set_error("learned preconditioner: the network takes 16 flux moments (P3)");
    return LB_E_ARG;
  }
  LB_TRY(ensure_elem_h(op));
  const int64_t nphi = op->nmom * op->NENK;
  double rr = 0.0;
  LB_TRY(dot(op, nphi, r, r, &rr));
  const double rms = std::sqrt(rr / (double)nphi);
  *F = feat_base(op, r, g0);
  F->inv_s = rms > 0.0 ? 1.0 / rms : 0.0;
  return LB_OK;
}

// {1 / rms, output scale} from the squared norm rr of n values, with the
// arithmetic of feat_args() and mlp_apply() on the host (IEEE square root and
// division on both sides, so the two entry points get the same bits).
__global__ void rms_scale_kernel(const double* __restrict__ rr, int64_t n, double* __restrict__ out) {
  if (threadIdx.x == 0) {
    const double rms = sqrt(rr[0] / (double)n);
    const double inv_s = rms > 0.0 ? 1.0 / rms : 0.0;
    out[0] = inv_s;
    out[1] = inv_s > 0.0 ? 1.0 / inv_s : 0.0;
  }
}

// Scratch of the allocation-free apply: the dot product's partial sums, its
// result and the two scales.  Its size does not depend on the mesh.  Each GPU
// holds one copy for the life of the process; calls on one GPU share it.
constexpr int kScratchDoubles = kDotPartials + 8;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of `scratch` on each GPU, looked up once per device with that
// device current (the host-side name is only a stand-in for it).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_ptr(double** out) {
  int dev = 0;
  LB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("lb_mlp_apply_noalloc_dev: device ordinal out of range");
    return LB_E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    LB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return LB_OK;
}

}  // namespace

int mlp_set(lb_op* op, const float* params, int nparams, int fmt) {
  if (!params || nparams != kParams || (fmt != LB_FMT_BF16 && fmt != LB_FMT_FP32)) {
    set_error("lb_op_set_preconditioner: expects LB_MLP_PARAMS FP32 values and LB_FMT_BF16 or LB_FMT_FP32");
    return LB_E_ARG;
  }
  if (op->nmom != 16) {
    set_error("lb_op_set_preconditioner: the network takes 16 flux moments (P3 quadrature sets only)");
    return LB_E_ARG;
  }
  for (int i = 0; i < nparams; ++i)
    if (!std::isfinite(params[i])) {
      set_error("lb_op_set_preconditioner: non-finite parameter");
      return LB_E_ARG;
    }
  std::vector<uint16_t> pk(kPackLen, 0);
  auto pack = [&](int layer, int off, int wofs, int K, int J) {
    const int tiles = J / 16, steps = (layer == 1) ? 1 : 2;
    for (int t = 0; t < tiles; ++t)
      for (int st = 0; st < steps; ++st)
        for (int lane = 0; lane < 64; ++lane)
          for (int q = 0; q < 8; ++q) {
            const int in = slot_input(layer, lane >> 4, st, q);
            const int out = 16 * t + (lane & 15);
            const float v = in < K ? params[wofs + in * J + out] : 0.f;
            pk[off + ((t * steps + st) * 64 + lane) * 8 + q] = bf16_rne(v);
          }
  };
  pack(1, kP1, kW1, kIn, kH);
  pack(2, kP2, kW2, kH, kH);
  pack(3, kP3, kW3, kH, kH);
  pack(4, kP4, kW4, kH, kOut);
  LB_TRY(op->d_mlp_w.reserve(kParams * sizeof(float)));
  LB_HIP(hipMemcpy(op->d_mlp_w.p, params, kParams * sizeof(float), hipMemcpyHostToDevice));
  LB_TRY(op->d_mlp_pack.reserve(kPackLen * sizeof(uint16_t)));
  LB_HIP(hipMemcpy(op->d_mlp_pack.p, pk.data(), kPackLen * sizeof(uint16_t), hipMemcpyHostToDevice));
  LB_TRY(ensure_elem_h(op));   // so that applying the network needs no allocation
  op->mlp = true;
  op->mlp_fmt = fmt;
  return LB_OK;
}

int mlp_features(lb_op* op, const double* r, int g0, float* feat, double* scale) {
  FeatArgs F;
  LB_TRY(feat_args(op, r, g0, &F));
  hipLaunchKernelGGL(features_kernel, dim3(blocks(op->NENK)), dim3(256), 0, op->stream, F, feat);
  LB_HIP(hipGetLastError());
  if (scale) *scale = F.inv_s > 0.0 ? 1.0 / F.inv_s : 0.0;
  return LB_OK;
}

// The network's kernel for the operator's format, scales from F (host) or F.dev_scale.
int launch_mlp(lb_op* op, const FeatArgs& F, double s, double* z) {
  if (op->mlp_fmt == LB_FMT_FP32) {
    hipLaunchKernelGGL(mlp_fp32_kernel, dim3(blocks(op->NENK)), dim3(256), 0, op->stream, F,
                       op->d_mlp_w.as<const float>(), s, z);
  } else {
    const int64_t tiles = (op->NENK + 15) / 16;
    const unsigned grid = (unsigned)std::min<int64_t>((tiles + 3) / 4, 8192);
    hipLaunchKernelGGL(mlp_bf16_kernel, dim3(grid), dim3(256), 0, op->stream, F, op->d_mlp_pack.as<const __bf16>(),
                       op->d_mlp_w.as<const float>(), s, z);
  }
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int mlp_apply(lb_op* op, const double* r, double* z, int g0) {
  if (!op->mlp) {
    set_error("learned preconditioner: no network set (lb_op_set_preconditioner)");
    return LB_E_NOT_READY;
  }
  FeatArgs F;
  LB_TRY(feat_args(op, r, g0, &F));
  const double s = F.inv_s > 0.0 ? 1.0 / F.inv_s : 0.0;
  return launch_mlp(op, F, s, z);
}

// mlp_apply() without allocation or host synchronisation: the residual's norm
// and the scales live in `scratch` (every entry is written before it is read,
// so nothing needs clearing), and the kernels read the scales from there.
int mlp_apply_noalloc(lb_op* op, const double* r, double* z, int g0) {
  if (!op->mlp || !op->d_elem_h.p) {
    set_error("lb_mlp_apply_noalloc_dev: no network set (lb_op_set_preconditioner)");
    return LB_E_NOT_READY;
  }
  if (op->nmom != 16) {
    set_error("lb_mlp_apply_noalloc_dev: the network takes 16 flux moments (P3)");
    return LB_E_ARG;
  }
  static_assert(kDotPartials + 3 <= kScratchDoubles, "scratch holds the partial sums, the norm and two scales");
  double* sp = nullptr;
  LB_TRY(scratch_ptr(&sp));
  const int64_t nphi = op->nmom * op->NENK;
  LB_TRY(launch_dot(op->stream, nphi, r, r, sp, sp + kDotPartials));
  hipLaunchKernelGGL(rms_scale_kernel, dim3(1), dim3(64), 0, op->stream, sp + kDotPartials, nphi,
                     sp + kDotPartials + 1);
  LB_HIP(hipGetLastError());
  FeatArgs F = feat_base(op, r, g0);
  F.dev_scale = sp + kDotPartials + 1;
  return launch_mlp(op, F, 0.0, z);
}

}  // namespace lb

using namespace lb;

extern "C" {

// Minibatch Adam on (X [N][18], Y [N][16]) device rows; params in/out on the host.
// At most 4096 rows per batch, so the loss partials (B * 16 / 256 blocks) fit the
// 256-entry buffer.
int lb_mlp_train_dev(const float* d_X, const float* d_Y, int64_t N, float* params, int nparams, int steps, int batch,
                     float lr, uint64_t seed, float* loss_hist) {
  return guarded([&]() -> int {
    if (!d_X || !d_Y || !params || nparams != kParams || N < 1 || N > INT32_MAX || steps < 1 || batch < 1 ||
        batch > 4096 || !(lr > 0.f)) {
      set_error("lb_mlp_train_dev: invalid argument");
      return LB_E_ARG;
    }
    hipStream_t st = nullptr;
    LB_HIP(hipStreamCreate(&st));
    const int B = batch;
    // every buffer in one allocation: params, grads, Adam m and v, step buffers, loss, indices
    const size_t nf = (size_t)4 * kParams + step_floats(B) + steps;
    const size_t bytes = nf * sizeof(float) + (size_t)steps * B * sizeof(int32_t);
    void* mem = nullptr;
    hipError_t e = hipMalloc(&mem, bytes);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(st);
      set_error(std::string("lb_mlp_train_dev: hipMalloc failed: ") + hipGetErrorString(e));
      return LB_E_HIP;
    }
    float* w = static_cast<float*>(mem);
    float* g = w + kParams;
    float* am = g + kParams;
    float* av = am + kParams;
    const StepBufs t = carve(av + kParams, B);
    float* loss = t.gpart;
    int32_t* idx = reinterpret_cast<int32_t*>(loss + steps);
    int rc = LB_OK;
    auto fail = [&](hipError_t err, const char* what) {
      set_error(std::string("lb_mlp_train_dev: ") + what + ": " + hipGetErrorString(err));
      rc = LB_E_HIP;
    };
    // minibatch rows: SplitMix64 stream `seed`
    std::vector<int32_t> h_idx((size_t)steps * B);
    for (size_t j = 0; j < h_idx.size(); ++j) {
      uint64_t zz = seed + (uint64_t)(j + 1) * 0x9E3779B97F4A7C15ull;
      zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
      zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
      zz = zz ^ (zz >> 31);
      h_idx[j] = (int32_t)(zz % (uint64_t)N);
    }
    if ((e = hipMemcpyAsync(idx, h_idx.data(), h_idx.size() * sizeof(int32_t), hipMemcpyHostToDevice, st)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(w, params, kParams * sizeof(float), hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemsetAsync(am, 0, 2 * kParams * sizeof(float), st)) != hipSuccess)
      fail(e, "upload");
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float bc1 = 1.f, bc2 = 1.f;
    for (int s = 0; s < steps && rc == LB_OK; ++s) {
      const int32_t* ix = idx + (size_t)s * B;
      enqueue_step(st, B, d_X, d_Y, ix, w, g, loss + s, t);
      bc1 *= b1;
      bc2 *= b2;
      hipLaunchKernelGGL(adam_kernel, dim3(blocks(kParams)), dim3(256), 0, st, kParams, w, g, am, av, lr, b1, b2, eps,
                         1.f - bc1, 1.f - bc2);
      if ((e = hipGetLastError()) != hipSuccess) fail(e, "launch");
    }
    if (rc == LB_OK) {
      if ((e = hipMemcpyAsync(params, w, kParams * sizeof(float), hipMemcpyDeviceToHost, st)) != hipSuccess)
        fail(e, "download");
      else if (loss_hist &&
               (e = hipMemcpyAsync(loss_hist, loss, steps * sizeof(float), hipMemcpyDeviceToHost, st)) != hipSuccess)
        fail(e, "download");
      else if ((e = hipStreamSynchronize(st)) != hipSuccess)
        fail(e, "synchronize");
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(mem);
    (void)hipStreamDestroy(st);
    return rc;
  });
}

// One gradient of the mean squared error over rows 0..B-1 (no update), for tests.
int lb_mlp_grad_dev(const float* d_X, const float* d_Y, int B, const float* params, int nparams, float* grad,
                    float* loss_out) {
  return guarded([&]() -> int {
    if (!d_X || !d_Y || !params || !grad || nparams != kParams || B < 1 || B > 4096) {
      set_error("lb_mlp_grad_dev: invalid argument");
      return LB_E_ARG;
    }
    void* mem = nullptr;
    const size_t nf = (size_t)2 * kParams + step_floats(B) + 1;
    LB_HIP(hipMalloc(&mem, nf * sizeof(float) + (size_t)B * sizeof(int32_t)));
    float* w = static_cast<float*>(mem);
    float* g = w + kParams;
    const StepBufs t = carve(g + kParams, B);
    float* loss = t.gpart;
    int32_t* idx = reinterpret_cast<int32_t*>(loss + 1);
    std::vector<int32_t> h_idx(B);
    for (int b = 0; b < B; ++b) h_idx[b] = b;
    hipStream_t st = nullptr;
    hipError_t e = hipSuccess;
    int rc = LB_OK;
    if ((e = hipMemcpy(idx, h_idx.data(), B * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(w, params, kParams * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
      set_error(std::string("lb_mlp_grad_dev: ") + hipGetErrorString(e));
      rc = LB_E_HIP;
    }
    if (rc == LB_OK) {
      enqueue_step(st, B, d_X, d_Y, idx, w, g, loss, t);
      if ((e = hipGetLastError()) != hipSuccess || (e = hipDeviceSynchronize()) != hipSuccess ||
          (e = hipMemcpy(grad, g, kParams * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess ||
          (loss_out && (e = hipMemcpy(loss_out, loss, sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)) {
        set_error(std::string("lb_mlp_grad_dev: ") + hipGetErrorString(e));
        rc = LB_E_HIP;
      }
    }
    (void)hipDeviceSynchronize();
    (void)hipFree(mem);
    return rc;
  });
}

}  // extern "C"
