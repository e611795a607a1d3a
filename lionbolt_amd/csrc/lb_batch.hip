// Batches of independent transport problems in one operator.
//
// The problems' meshes are concatenated into one mesh with no face between
// them (consecutive element ranges), so geometry, schedules, cross sections
// (one material per problem) and the moment products are those of one
// operator.  What is specific to a batch:
//   * the sweep order inside each direction is made problem-major (problem, level,
//     element), and every (problem, direction) pair becomes one work item of the
//     persistent sweep kernel, largest problems first;
//   * per-problem reductions and updates over [rows][NENK] vectors (segmented dot
//     products, axpy and scaling with one coefficient per problem);
//   * a moment-space GMRES(m) that runs the problems in lockstep, each with its own
//     Krylov basis coefficients, Hessenberg matrix, Givens rotations and stopping
//     test; one sweep launch per Arnoldi step serves every problem still iterating,
//     and converged problems drop out of the sweep.  Each problem's iterates are
//     those of an independent GMRES(m) (SPARSKIT's algorithm as in lb_krylov.cpp:
//     iters.f:1901-2206, mgsro's 0.98 re-orthogonalisation, Givens rotations).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "lb_internal.hpp"

namespace lb {

namespace {

constexpr int kSegBlock = 256;
constexpr int64_t kTileLen = 16384;

__device__ __forceinline__ double wg_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = kSegBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// One workgroup per tile: part[tile] = sum over the tile of x*y.
__global__ __launch_bounds__(kSegBlock) void seg_dot_partial_kernel(const int64_t* __restrict__ tiles,
                                                                    const double* __restrict__ x,
                                                                    const double* __restrict__ y,
                                                                    double* __restrict__ part) {
  __shared__ double sh[kSegBlock];
  const int64_t* t = tiles + 4 * (int64_t)blockIdx.x;
  const int64_t b = t[1], len = t[2];
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < len; j += kSegBlock) acc += x[b + j] * y[b + j];
  const double s = wg_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[p] = sum of problem p's tile partials, in tile order (deterministic).
__global__ void seg_dot_final_kernel(int P, const int32_t* __restrict__ tile_off, const double* __restrict__ part,
                                     double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double acc = 0.0;
  for (int j = tile_off[p]; j < tile_off[p + 1]; ++j) acc += part[j];
  out[p] = acc;
}

// y += coef[p] x over problem p's tiles (coef = 0 leaves a problem untouched).
__global__ __launch_bounds__(kSegBlock) void seg_axpy_kernel(const int64_t* __restrict__ tiles,
                                                             const double* __restrict__ coef,
                                                             const double* __restrict__ x, double* __restrict__ y) {
  const int64_t* t = tiles + 4 * (int64_t)blockIdx.x;
  const double a = coef[t[0]];
  if (a == 0.0) return;
  const int64_t b = t[1], len = t[2];
  for (int64_t j = threadIdx.x; j < len; j += kSegBlock) y[b + j] += a * x[b + j];
}

__global__ __launch_bounds__(kSegBlock) void seg_scale_kernel(const int64_t* __restrict__ tiles,
                                                              const double* __restrict__ coef, double* __restrict__ x) {
  const int64_t* t = tiles + 4 * (int64_t)blockIdx.x;
  const double a = coef[t[0]];
  if (a == 1.0) return;
  const int64_t b = t[1], len = t[2];
  for (int64_t j = threadIdx.x; j < len; j += kSegBlock) x[b + j] *= a;
}

}  // namespace

// Declares the operator's mesh as P independent problems in element ranges
// [eoff[p], eoff[p+1]) and builds the persistent sweep's item table.
int set_problems(lb_op* op, int P, const int64_t* eoff) {
  const int64_t NE = op->NE, NI = op->NI;
  if (P < 1 || !eoff || eoff[0] != 0 || eoff[P] != NE) {
    set_error("lb_op_set_problems: offsets must run from 0 to NE");
    return LB_E_ARG;
  }
  for (int p = 0; p < P; ++p)
    if (eoff[p + 1] <= eoff[p]) {
      set_error("lb_op_set_problems: every problem needs at least one element");
      return LB_E_ARG;
    }
  if (P > (1 << 20)) {
    set_error("lb_op_set_problems: too many problems");
    return LB_E_ARG;
  }
  std::vector<int32_t> prob(NE);
  for (int p = 0; p < P; ++p)
    for (int64_t e = eoff[p]; e < eoff[p + 1]; ++e) prob[e] = p;
  for (int64_t e = 0; e < NE; ++e)
    for (int f = 0; f < 4; ++f) {
      const int32_t ep = op->h_nbr[e * 4 + f];
      if (ep >= 0 && prob[ep] != prob[e]) {
        set_error("lb_op_set_problems: elements " + std::to_string(e) + " and " + std::to_string(ep) +
                  " share a face but lie in different problems");
        return LB_E_ARG;
      }
    }
  if (op->stream) LB_HIP(hipStreamSynchronize(op->stream));
  clear_graphs(op);
  BatchTable& B = op->batch;
  B.d_rec.release();
  B.d_rec_pos.release();
  B.d_inv.release();
  op->btune_pos.clear();
  for (auto& kv : B.seg) {
    kv.second.d_tiles.release();
    kv.second.d_tile_off.release();
  }
  B.seg.clear();
  B.P = P;
  B.eoff.assign(eoff, eoff + P + 1);
  B.pmat.resize(P);
  for (int p = 0; p < P; ++p) B.pmat[p] = op->h_mat[eoff[p]];
  // problem-major order inside each direction; the union's order is (level,
  // element) and every problem's levels start at 0, so a stable bucket by problem
  // gives (problem, level, element)
  const int ML = op->maxlev;
  B.tasks.assign((size_t)NI * NE, 0u);
  B.lvl.clear();
  std::vector<int32_t> items;
  std::vector<int64_t> fill(P);
  std::vector<int32_t> cur_lvl(P), base(P);
  for (int64_t i = 0; i < NI; ++i) {
    for (int p = 0; p < P; ++p) fill[p] = eoff[p];
    const int32_t* lo = &op->h_lvl_off[i * (ML + 1)];
    std::vector<std::vector<int32_t>> starts(P);
    for (int l = 0; l < op->h_nlev[i]; ++l)
      for (int32_t s = lo[l]; s < lo[l + 1]; ++s) {
        const uint32_t w = op->h_tasks[i * NE + s];
        const int p = prob[w & 0xFFFFFFu];
        if (starts[p].size() < (size_t)l + 1) {
          if (starts[p].size() != (size_t)l) {
            set_error("lb_op_set_problems: a problem's levels are not contiguous");
            return LB_E_ARG;
          }
          starts[p].push_back((int32_t)fill[p]);
        }
        B.tasks[i * NE + fill[p]++] = w;
      }
    for (int p = 0; p < P; ++p) {
      items.insert(items.end(), {p, (int32_t)i, (int32_t)B.lvl.size(), (int32_t)starts[p].size()});
      B.lvl.insert(B.lvl.end(), starts[p].begin(), starts[p].end());
      B.lvl.push_back((int32_t)eoff[p + 1]);
    }
  }
  // largest problems first, then longest level chains
  const int nitems = (int)(items.size() / 4);
  std::vector<int> ord(nitems);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    const int64_t na = eoff[items[4 * a] + 1] - eoff[items[4 * a]], nb = eoff[items[4 * b] + 1] - eoff[items[4 * b]];
    if (na != nb) return na > nb;
    return items[4 * a + 3] > items[4 * b + 3];
  });
  B.items.resize(items.size());
  for (int k = 0; k < nitems; ++k)
    for (int j = 0; j < 4; ++j) B.items[4 * k + j] = items[4 * ord[k] + j];
  LB_TRY(upload(op, B.d_tasks, B.tasks.data(), B.tasks.size() * sizeof(uint32_t)));
  LB_TRY(upload(op, B.d_items, B.items.data(), B.items.size() * sizeof(int32_t)));
  LB_TRY(upload(op, B.d_lvl, B.lvl.data(), B.lvl.size() * sizeof(int32_t)));
  std::vector<int32_t> ones(P, 1);
  LB_TRY(upload(op, B.d_active, ones.data(), ones.size() * sizeof(int32_t)));
  LB_TRY(B.d_counter.reserve(256));
  LB_TRY(B.d_coef.reserve((size_t)P * sizeof(double)));
  LB_TRY(B.d_out.reserve((size_t)P * sizeof(double)));
  if (B.h_io) (void)hipHostFree(B.h_io);
  B.h_io = nullptr;
  LB_HIP(hipHostMalloc((void**)&B.h_io, (size_t)P * sizeof(double), hipHostMallocDefault));
  int dev = 0;
  LB_HIP(hipGetDevice(&dev));
  LB_HIP(hipDeviceGetAttribute(&B.ncu, hipDeviceAttributeMultiprocessorCount, dev));
  op->btune.clear();
  return LB_OK;
}

SegTable* seg_table(lb_op* op, int rows, int* rc) {
  BatchTable& B = op->batch;
  auto it = B.seg.find(rows);
  if (it != B.seg.end()) return &it->second;
  SegTable t;
  t.tile_off.push_back(0);
  for (int p = 0; p < B.P; ++p) {
    const int64_t c0 = 4 * B.eoff[p], c1 = 4 * B.eoff[p + 1];
    for (int r = 0; r < rows; ++r)
      for (int64_t c = c0; c < c1; c += kTileLen)
        t.tiles.insert(t.tiles.end(), {p, r * op->NENK + c, std::min(kTileLen, c1 - c), 0});
    t.tile_off.push_back((int32_t)(t.tiles.size() / 4));
  }
  if ((*rc = upload(op, t.d_tiles, t.tiles.data(), t.tiles.size() * sizeof(int64_t))) != LB_OK) {
    t.d_tiles.release();
    return nullptr;
  }
  if ((*rc = upload(op, t.d_tile_off, t.tile_off.data(), t.tile_off.size() * sizeof(int32_t))) != LB_OK) {
    t.d_tiles.release();
    t.d_tile_off.release();
    return nullptr;
  }
  return &B.seg.emplace(rows, std::move(t)).first->second;
}

// out[p] (host) = x_p . y_p for every problem.
int seg_dot(lb_op* op, SegTable* t, const double* x, const double* y, double* out) {
  BatchTable& B = op->batch;
  const int ntiles = (int)(t->tiles.size() / 4);
  LB_TRY(B.d_part.reserve((size_t)ntiles * sizeof(double)));
  hipLaunchKernelGGL(seg_dot_partial_kernel, dim3(ntiles), dim3(kSegBlock), 0, op->stream,
                     t->d_tiles.as<const int64_t>(), x, y, B.d_part.as<double>());
  hipLaunchKernelGGL(seg_dot_final_kernel, dim3((B.P + 63) / 64), dim3(64), 0, op->stream, B.P,
                     t->d_tile_off.as<const int32_t>(), B.d_part.as<const double>(), B.d_out.as<double>());
  LB_HIP(hipGetLastError());
  LB_HIP(hipMemcpyAsync(B.h_io, B.d_out.p, (size_t)B.P * sizeof(double), hipMemcpyDeviceToHost, op->stream));
  LB_HIP(hipStreamSynchronize(op->stream));
  std::copy(B.h_io, B.h_io + B.P, out);
  return LB_OK;
}

static int upload_coef(lb_op* op, const double* coef) {
  BatchTable& B = op->batch;
  // the staging buffer is reused by the next call, so the copy completes first
  std::copy(coef, coef + B.P, B.h_io);
  LB_HIP(hipMemcpyAsync(B.d_coef.p, B.h_io, (size_t)B.P * sizeof(double), hipMemcpyHostToDevice, op->stream));
  LB_HIP(hipStreamSynchronize(op->stream));
  return LB_OK;
}

int seg_axpy(lb_op* op, SegTable* t, const double* coef, const double* x, double* y) {
  LB_TRY(upload_coef(op, coef));
  hipLaunchKernelGGL(seg_axpy_kernel, dim3((unsigned)(t->tiles.size() / 4)), dim3(kSegBlock), 0, op->stream,
                     t->d_tiles.as<const int64_t>(), op->batch.d_coef.as<const double>(), x, y);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int seg_scale(lb_op* op, SegTable* t, const double* coef, double* x) {
  LB_TRY(upload_coef(op, coef));
  hipLaunchKernelGGL(seg_scale_kernel, dim3((unsigned)(t->tiles.size() / 4)), dim3(kSegBlock), 0, op->stream,
                     t->d_tiles.as<const int64_t>(), op->batch.d_coef.as<const double>(), x);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

static int set_active(lb_op* op, const std::vector<char>& act) {
  BatchTable& B = op->batch;
  std::vector<int32_t> a(B.P);
  for (int p = 0; p < B.P; ++p) a[p] = act[p] ? 1 : 0;
  LB_HIP(hipMemcpyAsync(B.d_active.p, a.data(), a.size() * sizeof(int32_t), hipMemcpyHostToDevice, op->stream));
  LB_HIP(hipStreamSynchronize(op->stream));
  return LB_OK;
}

// Lockstep moment-space GMRES(m) over the problems of a batch (single group g0 =
// 0).  qmom [nmom][NENK]: moments of each problem's external source (the
// discrete source is M qmom); phi [nmom][NENK]: collided flux moments out.
int batch_solve_dev(lb_op* op, const double* qmom, const lb_solver_opts* o, double* phi_out, lb_solve_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  BatchTable& B = op->batch;
  const int P = B.P;
  const int64_t n = op->NENK * op->NI, nphi = op->NENK * op->nmom;
  lb_solver_opts opt = o ? *o : lb_solver_opts{LB_SOLVER_GMRES, LB_SPACE_MOMENT, 30, 0, 1e-8, 1e-12, 0};
  const int m = opt.restart > 0 ? opt.restart : 30;
  const int maxit = opt.maxit > 0 ? opt.maxit : 100000;
  hipStream_t str = op->stream;
  int rc = LB_OK;
  SegTable* tm = seg_table(op, op->nmom, &rc);
  if (!tm) return rc;
  SegTable* tp = seg_table(op, (int)op->NI, &rc);
  if (!tp) return rc;
  LB_TRY(op->w_krylov.reserve((size_t)(m + 3) * nphi * sizeof(double)));
  LB_TRY(op->w_mg.reserve(n * sizeof(double)));
  LB_TRY(op->w_a2.reserve(n * sizeof(double)));
  LB_TRY(op->w_b.reserve(n * sizeof(double)));
  LB_TRY(op->w_phi.reserve(nphi * sizeof(double)));
  double* V = op->w_krylov.as<double>();
  double* rhs = V + (int64_t)(m + 1) * nphi;
  double* r = rhs + nphi;
  double* q = op->w_b.as<double>();       // discrete source M qmom
  double* psi_t = op->w_a2.as<double>();
  double* tphi = op->w_phi.as<double>();
  double* sol = phi_out;
  auto vec = [&](int j) { return V + (int64_t)j * nphi; };
  int sweeps = 0;
  std::vector<char> all(P, 1);
  LB_TRY(set_active(op, all));

  // rhs = D T^-1 q; with a sweep-order flux the persistent sweep writes it to w_so
  // in the batch's problem-major order and D gathers it back by element
  const bool so = sweep_order_flux(op);
  double* psi_so = nullptr;
  if (so) {
    LB_TRY(op->w_so.reserve(n * sizeof(double)));
    psi_so = op->w_so.as<double>();
  }
  LB_TRY(launch_moments_to_discrete(op, qmom, q, 0, 1, 0, false));
  // ||T^-1 q|| per problem, the true residual's denominator: a problem's positions
  // in the batch's sweep order are its element range, so the segments are the same
  std::vector<double> bb(P), rr(P);
  if (so) {
    LB_TRY(launch_sweep_pos(op, q, psi_so, 0));
    LB_TRY(seg_dot(op, tp, psi_so, psi_so, bb.data()));
    LB_TRY(launch_moments_pos(op, psi_so, rhs));
  } else {
    LB_TRY(launch_copy(str, n, q, psi_t));
    LB_TRY(launch_sweep(op, psi_t, psi_t, 0, 1));
    LB_TRY(launch_moments(op, psi_t, rhs, 1, op->NENK));
  }
  ++sweeps;
  // A v = v - D T^-1 M S v for every problem; problems not in `act` skip the sweep
  auto apply = [&](double* v) -> int {
    ++sweeps;
    LB_TRY(launch_moments_to_discrete(op, v, psi_t, 0, 1, 0));
    if (so) {
      LB_TRY(launch_sweep_pos(op, psi_t, psi_so, 0));
      LB_TRY(launch_moments_pos(op, psi_so, tphi));
    } else {
      LB_TRY(launch_sweep(op, psi_t, psi_t, 0, 1));
      LB_TRY(launch_moments(op, psi_t, tphi, 1, op->NENK));
    }
    return launch_axpby(str, nphi, 1.0, v, -1.0, tphi, v);
  };
  // x0 = rhs / sigma_t of each problem
  std::vector<double> coef(P), rn(P), eps(P), zero(P, 0.0);
  LB_TRY(launch_copy(str, nphi, rhs, sol));
  for (int p = 0; p < P; ++p) coef[p] = 1.0 / op->h_sigt[(size_t)B.pmat[p] * op->G + 0];
  LB_TRY(seg_scale(op, tm, coef.data(), sol));
  LB_TRY(seg_dot(op, tm, rhs, rhs, rn.data()));
  for (int p = 0; p < P; ++p) eps[p] = opt.rtol * std::sqrt(rn[p]) + opt.atol;

  std::vector<char> active(P, 1), cyc(P);
  std::vector<int> iters(P, 0), status(P, LB_E_MAXIT), kdim(P);
  std::vector<double> res(P, 0.0);
  std::vector<std::vector<double>> H(P, std::vector<double>((size_t)(m + 1) * m, 0.0)), c(P, std::vector<double>(m)),
      sn(P, std::vector<double>(m)), rs(P, std::vector<double>(m + 1));
  std::vector<double> hh(P), nrm0(P), thr(P), fct(P), nb(P);
  auto any = [](const std::vector<char>& f) { return std::any_of(f.begin(), f.end(), [](char x) { return x != 0; }); };
  int cycles = 0;
  while (any(active)) {
    LB_TRY(set_active(op, active));
    // r = rhs - A sol
    LB_TRY(launch_copy(str, nphi, sol, r));
    LB_TRY(apply(r));
    LB_TRY(launch_axpby(str, nphi, 1.0, rhs, -1.0, r, r));
    LB_TRY(seg_dot(op, tm, r, r, nb.data()));
    for (int p = 0; p < P; ++p) {
      cyc[p] = 0;
      if (!active[p]) continue;
      const double beta = std::sqrt(nb[p]);
      res[p] = beta;
      if (std::isnan(beta)) {
        status[p] = LB_E_NAN;
        active[p] = 0;
      } else if (beta <= eps[p]) {
        status[p] = 0;
        active[p] = 0;
      } else if (iters[p] >= maxit) {
        active[p] = 0;
      } else {
        cyc[p] = 1;
        rs[p].assign(m + 1, 0.0);
        rs[p][0] = beta;
        coef[p] = 1.0 / beta;
        kdim[p] = 0;
      }
    }
    if (!any(cyc)) break;
    for (int p = 0; p < P; ++p)
      if (!cyc[p]) coef[p] = 0.0;
    LB_TRY(launch_copy(str, nphi, r, vec(0)));
    {
      std::vector<double> sc(P);
      for (int p = 0; p < P; ++p) sc[p] = cyc[p] ? coef[p] : 1.0;
      LB_TRY(seg_scale(op, tm, sc.data(), vec(0)));
    }
    std::vector<char> step = cyc;  // problems still extending their basis in this cycle
    for (int k = 1; k <= m && any(step); ++k) {
      LB_TRY(set_active(op, step));
      double* vk = vec(k);
      LB_TRY(launch_copy(str, nphi, vec(k - 1), vk));
      LB_TRY(apply(vk));
      LB_TRY(seg_dot(op, tm, vk, vk, nrm0.data()));
      for (int p = 0; p < P; ++p) thr[p] = nrm0[p] * 0.98;
      // modified Gram-Schmidt with selective re-orthogonalisation, per problem
      for (int i = 0; i < k; ++i) {
        LB_TRY(seg_dot(op, tm, vk, vec(i), fct.data()));
        for (int p = 0; p < P; ++p) {
          hh[p] = step[p] ? fct[p] : 0.0;
          coef[p] = -hh[p];
        }
        LB_TRY(seg_axpy(op, tm, coef.data(), vec(i), vk));
        bool again = false;
        for (int p = 0; p < P; ++p) again |= step[p] && fct[p] * fct[p] > thr[p];
        if (again) {
          LB_TRY(seg_dot(op, tm, vk, vec(i), fct.data()));
          for (int p = 0; p < P; ++p) {
            const bool use = step[p] && hh[p] * hh[p] > thr[p];
            coef[p] = use ? -fct[p] : 0.0;
            if (use) hh[p] += fct[p];
          }
          LB_TRY(seg_axpy(op, tm, coef.data(), vec(i), vk));
        }
        for (int p = 0; p < P; ++p) {
          if (!step[p]) continue;
          H[p][(size_t)(k - 1) * (m + 1) + i] = hh[p];
          nrm0[p] = std::max(0.0, nrm0[p] - hh[p] * hh[p]);
          thr[p] = nrm0[p] * 0.98;
        }
      }
      LB_TRY(seg_dot(op, tm, vk, vk, nb.data()));
      for (int p = 0; p < P; ++p) {
        coef[p] = 1.0;
        if (!step[p]) continue;
        const double nrm1 = std::sqrt(nb[p]);
        double* col = &H[p][(size_t)(k - 1) * (m + 1)];
        col[k] = nrm1;
        if (nrm1 > 0.0) coef[p] = 1.0 / nrm1;
        // previous rotations, then a new one (Givens, iters.f:3225-3260)
        for (int i = 1; i <= k - 1; ++i) {
          const double a = col[i - 1];
          col[i - 1] = c[p][i - 1] * a + sn[p][i - 1] * col[i];
          col[i] = c[p][i - 1] * col[i] - sn[p][i - 1] * a;
        }
        double x = col[k - 1], y = col[k], cc, ss;
        if (x == 0.0 && y == 0.0) {
          cc = 1.0;
          ss = 0.0;
        } else if (std::fabs(y) > std::fabs(x)) {
          const double t = x / y;
          x = std::sqrt(1.0 + t * t);
          ss = std::copysign(1.0 / x, y);
          cc = t * ss;
        } else {
          const double t = y / x;
          y = std::sqrt(1.0 + t * t);
          cc = std::copysign(1.0 / y, x);
          ss = t * cc;
        }
        c[p][k - 1] = cc;
        sn[p][k - 1] = ss;
        col[k - 1] = cc * col[k - 1] + ss * col[k];
        col[k] = 0.0;
        rs[p][k] = -ss * rs[p][k - 1];
        rs[p][k - 1] = cc * rs[p][k - 1];
        kdim[p] = k;
        ++iters[p];
        res[p] = std::fabs(rs[p][k]);
        if (std::isnan(res[p])) {
          status[p] = LB_E_NAN;
          step[p] = 0;
          active[p] = 0;
        } else if (res[p] <= eps[p] || nrm1 <= 0.0 || iters[p] >= maxit) {
          step[p] = 0;
        }
      }
      LB_TRY(seg_scale(op, tm, coef.data(), vk));
    }
    // sol += V y with H y = rs (upper triangular), per problem
    std::vector<std::vector<double>> y(P);
    for (int p = 0; p < P; ++p) {
      if (!cyc[p] || status[p] == LB_E_NAN) continue;
      const int kk = kdim[p];
      y[p].assign(kk, 0.0);
      for (int i = kk - 1; i >= 0; --i) {
        double acc = rs[p][i];
        for (int j = i + 1; j < kk; ++j) acc -= H[p][(size_t)j * (m + 1) + i] * y[p][j];
        const double d = H[p][(size_t)i * (m + 1) + i];
        y[p][i] = d != 0.0 ? acc / d : 0.0;
      }
    }
    for (int i = 0; i < m; ++i) {
      bool use = false;
      for (int p = 0; p < P; ++p) {
        coef[p] = (i < (int)y[p].size()) ? y[p][i] : 0.0;
        use |= coef[p] != 0.0;
      }
      if (use) LB_TRY(seg_axpy(op, tm, coef.data(), vec(i), sol));
    }
    ++cycles;
  }
  // psi = T^-1 (M S phi + q); true transport residual per problem
  LB_TRY(set_active(op, all));
  if (so) {
    // T^-1 q - (I - T^-1 K) psi = T^-1 M S (D psi - phi), exact for this psi: two
    // sweep-order sweeps, with ||T^-1 q|| from the first sweep of the solve
    double* qs = op->w_mg.as<double>();
    LB_TRY(launch_copy(str, n, q, qs));
    LB_TRY(launch_moments_to_discrete(op, sol, qs, 0, 1, 1));
    LB_TRY(launch_sweep_pos(op, qs, psi_so, 0));
    LB_TRY(launch_moments_pos(op, psi_so, tphi));
    LB_TRY(launch_axpby(str, nphi, 1.0, tphi, -1.0, sol, tphi));
    LB_TRY(launch_moments_to_discrete(op, tphi, qs, 0, 1, 0));
    LB_TRY(launch_sweep_pos(op, qs, psi_so, 0));
    sweeps += 2;
    LB_TRY(seg_dot(op, tp, psi_so, psi_so, rr.data()));
  } else {
    double* psi = op->w_mg.as<double>();
    LB_TRY(launch_copy(str, n, q, psi));
    LB_TRY(launch_moments_to_discrete(op, sol, psi, 0, 1, 1));
    LB_TRY(launch_sweep(op, psi, psi, 0, 1));
    double* b = psi_t;
    LB_TRY(launch_copy(str, n, q, b));
    LB_TRY(launch_sweep(op, b, b, 0, 1));
    LB_TRY(launch_copy(str, n, psi, q));
    LB_TRY(boltzmann_matvec_dev(op, q, 0));
    sweeps += 3;
    LB_TRY(seg_dot(op, tp, b, b, bb.data()));
    LB_TRY(launch_axpby(str, n, 1.0, b, -1.0, q, q));
    LB_TRY(seg_dot(op, tp, q, q, rr.data()));
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  int rc_all = LB_OK;
  for (int p = 0; p < P; ++p) {
    if (stats) {
      lb_solve_stats& s = stats[p];
      s.status = status[p];
      s.iterates = iters[p];
      s.matvecs = iters[p] + cycles;
      s.sweeps = sweeps;
      s.final_norm = res[p];
      s.true_residual = bb[p] > 0 ? std::sqrt(rr[p] / bb[p]) : std::sqrt(rr[p]);
      s.seconds = secs;
    }
    if (status[p] != 0 && rc_all == LB_OK) rc_all = status[p];
  }
  if (rc_all != LB_OK)
    set_error(rc_all == LB_E_MAXIT ? "batched GMRES : did not converge in the iteration limit"
                                   : "batched GMRES : RESIDUAL NORM is NaN");
  return rc_all;
}

}  // namespace lb
