// Kernels of the group-batched multigroup solve: B consecutive energy groups
// g0 .. g0+B-1 solved as one coupled flux-moment system, so that every operator
// application is one B-group sweep.  Batch vectors use the sweep's group-fastest
// layout: moments [lm][sd * B + g], discrete [i][e][k][g].
#include <hip/hip_runtime.h>

#include "lb_internal.hpp"

namespace lb {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// acc[lm][sd*B+g] = qmom[g0+g][lm][sd] + sum_{g' < g0} sigma_{l(lm)}(g' -> g0+g, mat) phi[g'][lm][sd]:
// the external source plus the downscatter from the groups already solved
// (the EI source of Main.f90:584-629 in moment form).
__global__ __launch_bounds__(kBlock) void batch_source_kernel(const double* __restrict__ phi_all,
                                                              const double* __restrict__ S,
                                                              const double* __restrict__ qmom, int g0, int B, int G,
                                                              int Nm, int nmom, int L,
                                                              const int32_t* __restrict__ mat,
                                                              const int32_t* __restrict__ deg, int64_t NENK,
                                                              double* __restrict__ acc) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= (int64_t)nmom * NENK * B) return;
  const int64_t row = NENK * B;
  const int lm = int(j / row);
  const int64_t rem = j - (int64_t)lm * row;
  const int64_t sd = rem / B;
  const int g = g0 + int(rem - sd * B);
  const int m_el = mat[sd >> 2];
  const int l = deg[lm];
  const int64_t o = (int64_t)lm * NENK + sd;
  double a = qmom ? qmom[(int64_t)g * nmom * NENK + o] : 0.0;
  for (int gp = 0; gp < g0; ++gp)
    a += S[(((int64_t)gp * G + g) * Nm + m_el) * (L + 1) + l] * phi_all[(int64_t)gp * nmom * NENK + o];
  acc[j] = a;
}

// out[lm][sd*B+g] = sum_{g' <= g in the batch} sigma_{l(lm)}(g0+g' -> g0+g, mat) v[lm][sd*B+g']:
// within-group scattering and the downscatter inside the batch.
__global__ __launch_bounds__(kBlock) void batch_transfer_kernel(const double* __restrict__ v,
                                                                const double* __restrict__ S, int g0, int B, int G,
                                                                int Nm, int nmom, int L,
                                                                const int32_t* __restrict__ mat,
                                                                const int32_t* __restrict__ deg, int64_t NENK,
                                                                double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= (int64_t)nmom * NENK * B) return;
  const int64_t row = NENK * B;
  const int lm = int(j / row);
  const int64_t rem = j - (int64_t)lm * row;
  const int64_t sd = rem / B;
  const int g = int(rem - sd * B);
  const int m_el = mat[sd >> 2];
  const int l = deg[lm];
  const int64_t base = j - g;
  double a = 0.0;
  for (int gp = 0; gp <= g; ++gp)
    a += S[(((int64_t)(g0 + gp) * G + g0 + g) * Nm + m_el) * (L + 1) + l] * v[base + gp];
  out[j] = a;
}

// phi_all[g0+g][lm][sd] = Phi[lm][sd*B+g]
__global__ __launch_bounds__(kBlock) void batch_store_kernel(const double* __restrict__ Phi, int g0, int B, int nmom,
                                                             int64_t NENK, double* __restrict__ phi_all) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= (int64_t)nmom * NENK * B) return;
  const int64_t row = NENK * B;
  const int lm = int(j / row);
  const int64_t rem = j - (int64_t)lm * row;
  const int64_t sd = rem / B;
  const int g = int(rem - sd * B);
  phi_all[((int64_t)(g0 + g) * nmom + lm) * NENK + sd] = Phi[j];
}

// Phi[lm][sd*B+g] = phi_all[g0+g][lm][sd] (the inverse of batch_store_kernel)
__global__ __launch_bounds__(kBlock) void batch_gather_kernel(const double* __restrict__ phi_all, int g0, int B,
                                                              int nmom, int64_t NENK, double* __restrict__ Phi) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= (int64_t)nmom * NENK * B) return;
  const int64_t row = NENK * B;
  const int lm = int(j / row);
  const int64_t rem = j - (int64_t)lm * row;
  const int64_t sd = rem / B;
  const int g = int(rem - sd * B);
  Phi[j] = phi_all[((int64_t)(g0 + g) * nmom + lm) * NENK + sd];
}

struct GroupWeights {
  double w[kMaxGroupBatch];
};

// x[j] *= w[j % B] (every batch vector's innermost index is the group).
__global__ __launch_bounds__(kBlock) void group_scale_kernel(int64_t n, int B, GroupWeights W, double* __restrict__ x) {
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock)
    x[j] *= W.w[j % B];
}

constexpr int kGroupDotBlocks = 1024;

// part[block][g] = sum over this block's share of the j with j % B == g of x[j] y[j].
__global__ __launch_bounds__(kBlock) void group_dot_partial_kernel(int64_t n, int B, const double* __restrict__ x,
                                                                   const double* __restrict__ y,
                                                                   double* __restrict__ part) {
  __shared__ double sh[kMaxGroupBatch][kBlock];
  double acc[kMaxGroupBatch];
#pragma unroll
  for (int g = 0; g < kMaxGroupBatch; ++g) acc[g] = 0.0;
  // the group of element j is j % B; it advances by stride % B per step
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  int g = int(((int64_t)blockIdx.x * kBlock + threadIdx.x) % B);
  const int gstep = int(stride % B);
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
    const double v = x[j] * y[j];
#pragma unroll
    for (int q = 0; q < kMaxGroupBatch; ++q)
      if (q == g) acc[q] += v;   // static register index
    g += gstep;
    if (g >= B) g -= B;
  }
#pragma unroll
  for (int q = 0; q < kMaxGroupBatch; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int q = 0; q < kMaxGroupBatch; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if ((int)threadIdx.x < B) part[(int64_t)blockIdx.x * kMaxGroupBatch + threadIdx.x] = sh[threadIdx.x][0];
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

// The same for B = 8 and 16-byte aligned vectors: each lane takes whole rows of
// B values (one per group), four 16-byte loads per operand and row, so the
// group of every value is its position in the row.
template <int B>
__global__ __launch_bounds__(kBlock) void group_dot_rows_kernel(int64_t rows, const double* __restrict__ x,
                                                                const double* __restrict__ y, int same,
                                                                double* __restrict__ part) {
  __shared__ double sh[B][kBlock];
  double acc[B];
#pragma unroll
  for (int g = 0; g < B; ++g) acc[g] = 0.0;
  const dbl2* xv = reinterpret_cast<const dbl2*>(x);
  const dbl2* yv = reinterpret_cast<const dbl2*>(y);
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < rows; r += stride) {
    dbl2 a[B / 2], b[B / 2];
#pragma unroll
    for (int h = 0; h < B / 2; ++h) a[h] = __builtin_nontemporal_load(xv + r * (B / 2) + h);
    if (same) {
#pragma unroll
      for (int h = 0; h < B / 2; ++h) b[h] = a[h];
    } else {
#pragma unroll
      for (int h = 0; h < B / 2; ++h) b[h] = __builtin_nontemporal_load(yv + r * (B / 2) + h);
    }
#pragma unroll
    for (int h = 0; h < B / 2; ++h) {
      acc[2 * h] += a[h].x * b[h].x;
      acc[2 * h + 1] += a[h].y * b[h].y;
    }
  }
#pragma unroll
  for (int g = 0; g < B; ++g) sh[g][threadIdx.x] = acc[g];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int g = 0; g < B; ++g) sh[g][threadIdx.x] += sh[g][threadIdx.x + s];
    __syncthreads();
  }
  if ((int)threadIdx.x < B) part[(int64_t)blockIdx.x * kMaxGroupBatch + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void group_dot_final_kernel(int nblocks, int B, const double* __restrict__ part, double* __restrict__ out) {
  const int g = threadIdx.x;
  if (g >= B) return;
  double a = 0.0;
  for (int b = 0; b < nblocks; ++b) a += part[(int64_t)b * kMaxGroupBatch + g];
  out[g] = a;
}

}  // namespace

int launch_batch_source(lb_op* op, const double* phi_all, const double* S, const double* qmom, int g0, int B,
                        double* acc) {
  const int64_t tot = (int64_t)op->nmom * op->NENK * B;
  hipLaunchKernelGGL(batch_source_kernel, dim3(grid_for(tot)), dim3(kBlock), 0, op->stream, phi_all, S, qmom, g0, B,
                     op->G, op->Nm, op->nmom, op->L, op->d_mat.as<const int32_t>(), op->d_deg.as<const int32_t>(),
                     op->NENK, acc);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_batch_transfer(lb_op* op, const double* S, int g0, int B, const double* v, double* out) {
  const int64_t tot = (int64_t)op->nmom * op->NENK * B;
  hipLaunchKernelGGL(batch_transfer_kernel, dim3(grid_for(tot)), dim3(kBlock), 0, op->stream, v, S, g0, B, op->G,
                     op->Nm, op->nmom, op->L, op->d_mat.as<const int32_t>(), op->d_deg.as<const int32_t>(), op->NENK,
                     out);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_batch_store(lb_op* op, const double* Phi, int g0, int B, double* phi_all) {
  const int64_t tot = (int64_t)op->nmom * op->NENK * B;
  hipLaunchKernelGGL(batch_store_kernel, dim3(grid_for(tot)), dim3(kBlock), 0, op->stream, Phi, g0, B, (int)op->nmom,
                     op->NENK, phi_all);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_batch_gather(lb_op* op, const double* phi_all, int g0, int B, double* Phi) {
  const int64_t tot = (int64_t)op->nmom * op->NENK * B;
  hipLaunchKernelGGL(batch_gather_kernel, dim3(grid_for(tot)), dim3(kBlock), 0, op->stream, phi_all, g0, B,
                     (int)op->nmom, op->NENK, Phi);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_group_scale(lb_op* op, int64_t n, int B, const double* w, double* x) {
  if (B < 1 || B > kMaxGroupBatch) {
    set_error("group batch out of range");
    return LB_E_ARG;
  }
  GroupWeights W{};
  for (int g = 0; g < B; ++g) W.w[g] = w[g];
  const int64_t blocks = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(group_scale_kernel, dim3((unsigned)(blocks < 8192 ? (blocks > 0 ? blocks : 1) : 8192)),
                     dim3(kBlock), 0, op->stream, n, B, W, x);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

// out[g] (host) = sum over j with j % B == g of x[j] y[j]; fixed grid and
// reduction order, so the result is reproducible.
int group_dot(lb_op* op, int64_t n, int B, const double* x, const double* y, double* out) {
  if (B < 1 || B > kMaxGroupBatch) {
    set_error("group batch out of range");
    return LB_E_ARG;
  }
  LB_TRY(op->w_red.reserve(((size_t)kGroupDotBlocks * kMaxGroupBatch + kMaxGroupBatch) * sizeof(double)));
  if (!op->w_host.p || op->w_host.bytes < kMaxGroupBatch * sizeof(double)) {
    if (op->w_host.p) (void)hipHostFree(op->w_host.p);
    op->w_host.p = nullptr;
    LB_HIP(hipHostMalloc(&op->w_host.p, kMaxGroupBatch * sizeof(double), hipHostMallocDefault));
    op->w_host.bytes = kMaxGroupBatch * sizeof(double);
  }
  double* part = op->w_red.as<double>();
  double* res = part + (size_t)kGroupDotBlocks * kMaxGroupBatch;
  const bool rows8 = B == 8 && n % 8 == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(y) & 15) == 0;
  if (rows8)
    hipLaunchKernelGGL(group_dot_rows_kernel<8>, dim3(kGroupDotBlocks), dim3(kBlock), 0, op->stream, n / 8, x, y,
                       x == y ? 1 : 0, part);
  else
    hipLaunchKernelGGL(group_dot_partial_kernel, dim3(kGroupDotBlocks), dim3(kBlock), 0, op->stream, n, B, x, y,
                       part);
  hipLaunchKernelGGL(group_dot_final_kernel, dim3(1), dim3(64), 0, op->stream, kGroupDotBlocks, B, part, res);
  LB_HIP(hipGetLastError());
  LB_HIP(hipMemcpyAsync(op->w_host.p, res, B * sizeof(double), hipMemcpyDeviceToHost, op->stream));
  LB_HIP(hipStreamSynchronize(op->stream));
  for (int g = 0; g < B; ++g) out[g] = static_cast<double*>(op->w_host.p)[g];
  return LB_OK;
}

}  // namespace lb
