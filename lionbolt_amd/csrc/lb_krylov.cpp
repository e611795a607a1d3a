// Krylov drivers over device vectors.
//
// GMRES follows SPARSKIT's restarted gmres (third_party/SPARSKIT/iters.f:1901-2206)
// as GMRESWrapper drives it (src/solvers/SPARSKIT.f90:20-181): convergence test
// ipar(3) = 2, ||r|| <= rtol ||rhs|| + atol; modified Gram-Schmidt with selective
// re-orthogonalisation when fct^2 > 0.98 ||v||^2 (mgsro, iters.f:3465-3586); Givens
// rotations of Golub & van Loan (givens, iters.f:3225-3260); at every restart the
// residual is recomputed with one operator application.  The vectors and all
// O(n) work live on the GPU (FP64); the (m+1) x m Hessenberg recurrences run on
// the host in FP64, which is where SPARSKIT keeps them too.
//
// Source iteration follows SourceIteration.f90:3-82.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <functional>
#include <vector>

#include "lb_internal.hpp"

namespace lb {

namespace {

using Apply = std::function<int(double*)>;

struct GmresResult {
  int status = 0;       // 0, LB_E_MAXIT, LB_E_BREAKDOWN, LB_E_NAN
  int iterates = 0;     // reverse-communication returns of the reference
  int matvecs = 0;
  double final_norm = 0.0;
};

void print_row(int verbose, int iter, double v) {
  if (verbose) std::printf("             %3d       %24.16E\n", iter, v);
}

// Restarted GMRES(m) on device vectors of length n from the iterate in sol.
// V holds m+1 vectors.
int gmres(lb_op* op, int64_t n, const Apply& apply, const double* rhs, double* sol, double* V, int m, double rtol,
          double atol, int maxit, int verbose, GmresResult* res) {
  hipStream_t s = op->stream;
  std::vector<double> H((size_t)(m + 1) * (m + 2) / 2 + 2, 0.0), c(m + 2), sn(m + 2), rs(m + 3);
  auto hcol = [&](int k) { return &H[(size_t)k * (k - 1) / 2]; };
  auto vec = [&](int j) { return V + (int64_t)j * n; };
  int iter = 0, nmv = 0;
  double fpar4 = 0.0, fpar5 = 0.0;
  bool first = true;
  auto ret = [&](double v) {
    ++iter;
    print_row(verbose, iter, v);
    if (std::isnan(v)) return false;
    return true;
  };
  if (verbose) {
    std::printf("          -------------------------------------\n");
    std::printf("           ITERATE           RESIDUAL NORM     \n");
    std::printf("          ---------     -----------------------\n");
  }
  ret(0.0);
  int status = 0;
  while (true) {
    double* v0 = vec(0);
    LB_TRY(launch_copy(s, n, sol, v0));
    LB_TRY(apply(v0));
    ++nmv;
    LB_TRY(launch_axpby(s, n, 1.0, rhs, -1.0, v0, v0));
    double a2;
    LB_TRY(dot(op, n, v0, v0, &a2));
    double alpha = std::sqrt(a2);
    if (first) {
      double r2;
      LB_TRY(dot(op, n, rhs, rhs, &r2));
      fpar4 = rtol * std::sqrt(r2) + atol;
      first = false;
    }
    fpar5 = alpha;
    rs[1] = alpha;
    if (alpha <= fpar4) {
      if (!ret(fpar5)) status = LB_E_NAN;
      break;
    }
    if (std::isnan(alpha)) {
      ret(alpha);
      status = LB_E_NAN;
      break;
    }
    LB_TRY(launch_scale(s, n, 1.0 / alpha, v0));
    int k = 0, ierr = 0;
    while (true) {
      ++k;
      if (!ret(fpar5)) {
        status = LB_E_NAN;
        break;
      }
      double* vk = vec(k);
      LB_TRY(launch_copy(s, n, vec(k - 1), vk));
      LB_TRY(apply(vk));
      ++nmv;
      double* hh = hcol(k);
      double nrm0;
      LB_TRY(dot(op, n, vk, vk, &nrm0));
      double thr = nrm0 * 0.98;
      if (!(nrm0 > 0.0)) {
        ierr = -1;
        break;
      }
      for (int i = 0; i < k; ++i) {
        const double* vi = vec(i);
        double fct;
        LB_TRY(dot(op, n, vk, vi, &fct));
        hh[i] = fct;
        LB_TRY(launch_axpby(s, n, -fct, vi, 1.0, vk, vk));
        if (fct * fct > thr) {
          LB_TRY(dot(op, n, vk, vi, &fct));
          hh[i] = hh[i] + fct;
          LB_TRY(launch_axpby(s, n, -fct, vi, 1.0, vk, vk));
        }
        nrm0 = nrm0 - hh[i] * hh[i];
        if (nrm0 < 0.0) nrm0 = 0.0;
        thr = nrm0 * 0.98;
      }
      double n1;
      LB_TRY(dot(op, n, vk, vk, &n1));
      double nrm1 = std::sqrt(n1);
      hh[k] = nrm1;
      if (!(nrm1 > 0.0)) {
        ierr = -3;
        break;
      }
      LB_TRY(launch_scale(s, n, 1.0 / nrm1, vk));
      for (int i = 1; i <= k - 1; ++i) {
        double a = hh[i - 1];
        hh[i - 1] = c[i] * a + sn[i] * hh[i];
        hh[i] = c[i] * hh[i] - sn[i] * a;
      }
      double& x = hh[k - 1];
      double& y = hh[k];
      double cc, ss;
      if (x == 0.0 && y == 0.0) {
        cc = 1.0;
        ss = 0.0;
      } else if (std::fabs(y) > std::fabs(x)) {
        double t = x / y;
        x = std::sqrt(1.0 + t * t);
        ss = std::copysign(1.0 / x, y);
        cc = t * ss;
      } else if (std::fabs(y) <= std::fabs(x)) {
        double t = y / x;
        y = std::sqrt(1.0 + t * t);
        cc = std::copysign(1.0 / y, x);
        ss = t * cc;
      } else {
        x = 0.0;
        y = 0.0;
        cc = 1.0;
        ss = 0.0;
      }
      x = std::fabs(x * y);
      c[k] = cc;
      sn[k] = ss;
      double a2k = -ss * rs[k];
      rs[k] = cc * rs[k];
      rs[k + 1] = a2k;
      fpar5 = std::fabs(a2k);
      if (!(k < m && !(fpar5 <= fpar4) && (maxit <= 0 || nmv < maxit))) break;
    }
    if (status != 0) break;
    // triangular solve and update (iters.f:2100-2155)
    int kk = k;
    while (kk > 0 && hcol(kk)[kk - 1] == 0.0) --kk;
    if (kk > 0 && ierr == 0) {
      std::vector<double> y(rs.begin(), rs.end());
      y[kk] = y[kk] / hcol(kk)[kk - 1];
      for (int i = kk - 1; i >= 1; --i) {
        const double* col = hcol(i + 1);
        for (int ii = 1; ii <= i; ++ii) y[ii] = y[ii] - y[i + 1] * col[ii - 1];
        y[i] = y[i] / hcol(i)[i - 1];
      }
      double* u = vec(0);
      LB_TRY(launch_scale(s, n, y[1], u));
      for (int i = 1; i <= kk - 1; ++i) LB_TRY(launch_axpby(s, n, y[i + 1], vec(i), 1.0, u, u));
      LB_TRY(launch_axpby(s, n, 1.0, u, 1.0, sol, sol));
    }
    if (ierr != 0 || kk == 0) {
      ret(fpar5);
      status = LB_E_BREAKDOWN;
      break;
    }
    if ((maxit <= 0 || nmv < maxit) && fpar5 > fpar4) {
      if (!ret(fpar5)) {
        status = LB_E_NAN;
        break;
      }
      continue;
    }
    status = (fpar5 <= fpar4) ? 0 : LB_E_MAXIT;
    ret(fpar5);
    break;
  }
  res->status = status;
  res->iterates = iter;
  res->matvecs = nmv;
  res->final_norm = fpar5;
  if (verbose && status == 0) {
    std::printf("\n          CONVERGED.\n          Total number of iterates : %3d\n", iter);
    std::printf("          Final Residual Norm      : %24.16E\n\n", fpar5);
  }
  return LB_OK;
}

int sweep_inplace(lb_op* op, double* v, int g0) { return launch_sweep(op, v, v, g0, 1); }

// Givens rotation of Golub & van Loan as SPARSKIT's givens (iters.f:3225-3260):
// (x, y) -> (c, s) with c x + s y = r, -s x + c y = 0.
void givens(double x, double y, double* c, double* s) {
  if (x == 0.0 && y == 0.0) {
    *c = 1.0;
    *s = 0.0;
  } else if (std::fabs(y) > std::fabs(x)) {
    const double t = x / y;
    *s = std::copysign(1.0 / std::sqrt(1.0 + t * t), y);
    *c = t * *s;
  } else {
    const double t = y / x;
    *c = std::copysign(1.0 / std::sqrt(1.0 + t * t), x);
    *s = t * *c;
  }
}

// Orthogonalises w against the orthonormal vectors V_0 .. V_{c-1} (stride n) by
// classical Gram-Schmidt, repeated once when the first pass removes more than
// half of ||w||^2 (the criterion of Daniel, Gragg, Kaufman & Stewart, Math.
// Comp. 30 (1976) 772-795), and scales w to unit length.  h[i] receives the
// projection on V_i, *nu the norm of the orthogonalised w.  When the second pass
// again removes more than half, w lies in span V to working precision: *nu = 0
// and w is not scaled.  A non-finite dot gives *nu = NaN.  The norm is
// ||w - V h||^2 = ||w||^2 - ||h||^2, accurate to a few ulps whenever at most
// half of ||w||^2 is removed, so one pass is one read of V and w for the dots
// (one host synchronisation) and one read-write of w for the update and scaling.
int orthonormalize(lb_op* op, int64_t n, const double* V, int c, double* w, double* h, double* nu) {
  std::vector<double> d(c + 1);
  for (int pass = 0; pass < 2; ++pass) {
    LB_TRY(multi_dot(op, n, V, n, c, w, true, d.data()));
    double hh = 0.0;
    for (int i = 0; i < c; ++i) hh += d[i] * d[i];
    const double ww = d[c];
    if (!std::isfinite(ww) || !std::isfinite(hh)) {
      *nu = std::nan("");
      return LB_OK;
    }
    for (int i = 0; i < c; ++i) h[i] = pass == 0 ? d[i] : h[i] + d[i];
    const double rem = ww - hh;
    if (rem > 0.5 * ww) {
      *nu = std::sqrt(rem);
      return multi_axpy(op, n, V, n, c, d.data(), 1.0 / *nu, w);
    }
    if (pass == 0) LB_TRY(multi_axpy(op, n, V, n, c, d.data(), 1.0, w));
  }
  *nu = 0.0;
  return LB_OK;
}

using ApplyK = std::function<int(const double* in, double* out)>;

// Restarted GMRES(m) for (I - K) x = rhs from x0 = 0, with K applied out of
// place.  I - K and K have the same Krylov spaces, so the Arnoldi process runs
// on w = K v_j, whose orthogonalisation does not cancel the identity part:
// if K v_j = sum_{i<=j} h_ij v_i + nu v_{j+1}, then
// (I - K) v_j = sum_i (delta_ij - h_ij) v_i - nu v_{j+1}, the Hessenberg matrix
// of I - K.  Orthogonalisation as orthonormalize() above; the Hessenberg matrix,
// the Givens rotations, the residual estimate and the final triangular solve
// stay in device memory (arnoldi_step), so an Arnoldi step costs one host
// synchronisation, at its end, for the stopping test.  Iterate counting, the
// stopping rule ||r|| <= rtol ||rhs|| + atol and the explicit residual at every
// restart as gmres() above.  When the basis stops growing (nu = 0), the solution
// is updated and the explicit residual decides: converged, another cycle, or
// LB_E_BREAKDOWN if a cycle made no progress.
int gmres_shifted(lb_op* op, int64_t n, const ApplyK& applyK, const double* rhs, double* sol, double* V, int m,
                  double rtol, double atol, int maxit, int verbose, GmresResult* res) {
  hipStream_t s = op->stream;
  auto vec = [&](int j) { return V + (int64_t)j * n; };
  ArnoldiDev arn;
  LB_TRY(arnoldi_reserve(op, m, &arn));
  int iter = 0, nmv = 0, status = 0;
  auto ret = [&](double v) {
    ++iter;
    print_row(verbose, iter, v);
    return !std::isnan(v);
  };
  if (verbose) {
    std::printf("          -------------------------------------\n");
    std::printf("           ITERATE           RESIDUAL NORM     \n");
    std::printf("          ---------     -----------------------\n");
  }
  ret(0.0);
  LB_HIP(hipMemsetAsync(sol, 0, n * sizeof(double), s));
  double r2 = 0.0;
  LB_TRY(dot(op, n, rhs, rhs, &r2));
  const double eps = rtol * std::sqrt(r2) + atol;
  double resn = 0.0, last_beta = -1.0;
  bool initial = true;
  while (true) {
    double* v0 = vec(0);
    if (initial) {
      LB_TRY(launch_copy(s, n, rhs, v0));   // r0 = rhs - (I - K) 0
      initial = false;
    } else {                                // r = rhs - x + K x
      LB_TRY(applyK(sol, v0));
      ++nmv;
      LB_TRY(launch_axpby(s, n, 1.0, rhs, 1.0, v0, v0));
      LB_TRY(launch_axpby(s, n, -1.0, sol, 1.0, v0, v0));
    }
    double b2;
    LB_TRY(dot(op, n, v0, v0, &b2));
    const double beta = std::sqrt(b2);
    resn = beta;
    if (std::isnan(beta)) {
      ret(beta);
      status = LB_E_NAN;
      break;
    }
    if (beta <= eps) {
      ret(beta);
      break;
    }
    if (last_beta >= 0.0 && !(beta < last_beta)) {   // a cycle that ended early made no progress
      ret(beta);
      status = LB_E_BREAKDOWN;
      break;
    }
    if (maxit > 0 && nmv >= maxit) {
      ret(beta);
      status = LB_E_MAXIT;
      break;
    }
    LB_TRY(launch_scale(s, n, 1.0 / beta, v0));
    LB_TRY(arnoldi_start(op, arn, beta));
    int k = 0;
    bool invariant = false;
    while (k < m) {
      if (!ret(resn)) {
        status = LB_E_NAN;
        break;
      }
      LB_TRY(applyK(vec(k), vec(k + 1)));
      ++nmv;
      ArnoldiCtl st;
      LB_TRY(arnoldi_step(op, n, V, k, arn, true, &st));
      if (st.state == kArnNaN) {
        status = LB_E_NAN;
        break;
      }
      resn = st.resn;
      ++k;
      if (std::isnan(resn)) {
        status = LB_E_NAN;
        break;
      }
      if (st.state == kArnInvariant) {
        invariant = true;
        break;
      }
      if (resn <= eps || (maxit > 0 && nmv >= maxit)) break;
    }
    if (status != 0) break;
    // y = R^-1 g over the k columns, x += V y (one fused pass per 8 vectors)
    LB_TRY(arnoldi_update(op, n, V, k, arn, sol));
    if (invariant) {
      last_beta = beta;                // the explicit residual decides
      continue;
    }
    last_beta = -1.0;
    if (resn <= eps) {
      ret(resn);
      break;
    }
    if (maxit > 0 && nmv >= maxit) {
      ret(resn);
      status = LB_E_MAXIT;
      break;
    }
    if (!ret(resn)) {
      status = LB_E_NAN;
      break;
    }
  }
  res->status = status;
  res->iterates = iter;
  res->matvecs = nmv;
  res->final_norm = resn;
  if (verbose && status == 0) {
    std::printf("\n          CONVERGED.\n          Total number of iterates : %3d\n", iter);
    std::printf("          Final Residual Norm      : %24.16E\n\n", resn);
  }
  return LB_OK;
}

using Precond = std::function<int(const double*, double*)>;
using Snapshot = std::function<int(const double* sol, const double* Z, int k, const std::vector<double>& y)>;

struct FgmresResult {
  int status = 0, iterates = 0, matvecs = 0, fallback = 0;
  double final_norm = 0.0;
};

// Flexible GMRES(m) (Saad, SIAM J. Sci. Comput. 14 (1993) 461-469): any
// search direction z_k may enter the basis, A Z_k = V_{k+1} H_k still holds, and
// x = x0 + Z y minimises the residual over span(Z).  The learned preconditioner
// maps residuals to errors, so it is applied to the current residual r_k (the
// GMRESR / flexible GCR choice of van der Vorst & Vuik, 1994): z_k = P(r_k).
// r_k is carried along at the cost of one update per step,
// r_k = s_k^2 r_{k-1} + c_k rho_k v_k (rho_k the rotated right-hand side), and
// with P = I the basis is the Krylov space of plain GMRES.  Orthogonalisation by
// orthonormalize() above (classical Gram-Schmidt, once more when needed), Givens rotations,
// stopping rule ||r|| <= rtol ||rhs|| + atol and an explicit residual at every
// restart, as gmres() above.  Fallback, z_k = v_k (plain GMRES) for the rest of
// the solve: when any 3 consecutive preconditioned steps reduce the residual
// estimate by less than a factor 2 per step on average (a useful approximate
// inverse does better than that), or when P returns a non-finite or zero vector.
// A cycle that ends because the basis stopped growing leaves the decision to the
// next explicit residual, which must be smaller than the cycle's first one, else
// the solve ends with LB_E_BREAKDOWN (as gmres_shifted).
int fgmres(lb_op* op, int64_t n, const Apply& apply, const Precond* P, const double* rhs, double* sol, double* V,
           double* Z, double* R, int m, double rtol, double atol, int maxit, int verbose, const Snapshot* snap,
           FgmresResult* res, bool zero_x0 = false) {
  hipStream_t s = op->stream;
  auto vec = [&](int j) { return V + (int64_t)j * n; };
  auto zvec = [&](int j) { return Z + (int64_t)j * n; };
  std::vector<double> H((size_t)(m + 1) * m, 0.0), c(m), sn(m), g(m + 1), hk(m + 1), hist;
  auto h = [&](int i, int k) -> double& { return H[(size_t)k * (m + 1) + i]; };
  constexpr int kWindow = 3;
  constexpr double kDrop = 0.125;   // 0.5 per step over the window
  double beta0 = -1.0;
  double r2 = 0.0;
  LB_TRY(dot(op, n, rhs, rhs, &r2));
  const double eps = rtol * std::sqrt(r2) + atol;
  bool use_pc = P != nullptr;
  int iters = 0, nmv = 0, status = LB_E_MAXIT, fallback = 0;
  double resn = 0.0;
  auto solve_y = [&](int k, std::vector<double>& y) {
    y.assign(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
      double a = g[i];
      for (int j = i + 1; j < k; ++j) a -= h(i, j) * y[j];
      y[i] = h(i, i) != 0.0 ? a / h(i, i) : 0.0;
    }
  };
  bool done = false;
  double hn_last = 1.0;
  double last_beta = -1.0;   // explicit residual before a cycle that ended on an invariant space
  if (zero_x0) LB_HIP(hipMemsetAsync(sol, 0, n * sizeof(double), s));
  bool initial = zero_x0;   // x0 = 0: the first residual is rhs, no operator application
  while (!done) {
    double* v0 = vec(0);
    if (initial) {
      LB_TRY(launch_copy(s, n, rhs, v0));
      initial = false;
    } else {
      LB_TRY(launch_copy(s, n, sol, v0));
      LB_TRY(apply(v0));
      ++nmv;
      LB_TRY(launch_axpby(s, n, 1.0, rhs, -1.0, v0, v0));
    }
    double b2 = 0.0;
    LB_TRY(dot(op, n, v0, v0, &b2));
    const double beta = std::sqrt(b2);
    resn = beta;
    if (std::isnan(beta)) {
      status = LB_E_NAN;
      break;
    }
    if (beta <= eps) {
      status = 0;
      break;
    }
    if (last_beta >= 0.0 && !(beta < last_beta)) {   // a cycle that ended early made no progress
      status = LB_E_BREAKDOWN;
      break;
    }
    if (maxit > 0 && iters >= maxit) break;
    if (beta0 < 0.0) beta0 = beta;
    if (use_pc) LB_TRY(launch_copy(s, n, v0, R));   // r_0
    LB_TRY(launch_scale(s, n, 1.0 / beta, v0));
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    while (k < m) {
      double* zk = zvec(k);
      if (use_pc) {
        LB_TRY((*P)(R, zk));
        double zz = 0.0;
        LB_TRY(dot(op, n, zk, zk, &zz));
        if (!std::isfinite(zz) || zz == 0.0) {
          use_pc = false;
          fallback = iters > 0 ? iters : 1;
        }
      }
      if (!use_pc) LB_TRY(launch_copy(s, n, vec(k), zk));
      double* w = vec(k + 1);
      LB_TRY(launch_copy(s, n, zk, w));
      LB_TRY(apply(w));
      ++nmv;
      double hn = 0.0;
      LB_TRY(orthonormalize(op, n, V, k + 1, w, hk.data(), &hn));
      hn_last = hn;
      if (std::isnan(hn)) {
        status = LB_E_NAN;
        done = true;
        break;
      }
      for (int i = 0; i <= k; ++i) h(i, k) = hk[i];
      h(k + 1, k) = hn;
      for (int i = 0; i < k; ++i) {
        const double a = h(i, k), b = h(i + 1, k);
        h(i, k) = c[i] * a + sn[i] * b;
        h(i + 1, k) = c[i] * b - sn[i] * a;
      }
      givens(h(k, k), h(k + 1, k), &c[k], &sn[k]);
      h(k, k) = c[k] * h(k, k) + sn[k] * h(k + 1, k);
      h(k + 1, k) = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = c[k] * g[k];
      resn = std::fabs(g[k + 1]);
      if (use_pc && hn > 0.0)   // r_{k+1} = s^2 r_k + c rho_{k+1} v_{k+1}
        LB_TRY(launch_axpby(s, n, c[k] * g[k + 1], vec(k + 1), sn[k] * sn[k], R, R));
      ++k;
      ++iters;
      hist.push_back(resn);
      print_row(verbose, iters, resn);
      if (snap) {
        std::vector<double> y;
        solve_y(k, y);
        LB_TRY((*snap)(sol, Z, k, y));
      }
      if (std::isnan(resn)) {
        status = LB_E_NAN;
        done = true;
        break;
      }
      const double before = (int)hist.size() > kWindow ? hist[hist.size() - 1 - kWindow]
                            : ((int)hist.size() == kWindow ? beta0 : -1.0);
      if (use_pc && before > 0.0 && resn > kDrop * before) {
        use_pc = false;
        fallback = iters;
      }
      if (resn <= eps || !(hn > 0.0) || (maxit > 0 && iters >= maxit)) break;
    }
    if (status == LB_E_NAN) break;
    std::vector<double> y;
    solve_y(k, y);
    for (double& t : y) t = -t;
    LB_TRY(multi_axpy(op, n, Z, n, k, y.data(), 1.0, sol));   // x += Z y
    // A Z = V H holds for any z_k, so |g_k| is the residual norm of the update,
    // as in plain GMRES: a cycle that reached eps ends the solve without an
    // explicit residual (a basis that stopped growing still gets one).
    if (resn <= eps && hn_last > 0.0) {
      status = 0;
      done = true;
    }
    // a basis that stopped growing hands the decision to the explicit residual,
    // which must then be smaller than this cycle's starting residual
    last_beta = hn_last > 0.0 ? -1.0 : beta;
  }
  res->status = status;
  res->iterates = iters;
  res->matvecs = nmv;
  res->fallback = fallback;
  res->final_norm = resn;
  return LB_OK;
}

}  // namespace

double group_thickness(const lb_op* op, int g) {
  double vol = 0.0;
  for (double d : op->h_detJ) vol += d / 6.0;
  double st = 0.0;
  for (int m = 0; m < op->Nm; ++m) st = std::max(st, op->h_sigt[(size_t)m * op->G + g]);
  return st * std::cbrt(vol);
}

// out <- D T^-1 M S in: the scattering-sweep-moments operator K of group g0 on
// flux moments, out of place.  With a sweep-order flux the sweep writes its
// flux to w_so and D gathers it back by element; otherwise the sweep runs in
// place on M S in.
int moment_k_dev(lb_op* op, const double* in, double* out, int g0) {
  const int64_t n = op->NENK * op->NI;
  if (sweep_order_fp32(op)) {
    // the flux for D in FP32 (the sweep's arithmetic and its traces stay FP64)
    LB_TRY(op->w_a2.reserve(n * sizeof(double)));
    LB_TRY(op->w_so.reserve(n * sizeof(double)));
    LB_TRY(op->w_so32.reserve(n * sizeof(float)));
    LB_TRY(launch_moments_to_discrete(op, in, op->w_a2.as<double>(), g0, 1, 0));
    LB_TRY(launch_sweep_f32(op, op->w_a2.as<const double>(), op->w_so.as<double>(), op->w_so32.as<float>(), g0));
    return launch_moments_pos32(op, op->w_so32.as<const float>(), out);
  }
  if (sweep_order_flux(op)) {
    LB_TRY(op->w_so.reserve(n * sizeof(double)));
    double* so = op->w_so.as<double>();
    if (op->flux_order == LB_FLUX_SWEEP_ORDER_SOURCE) {
      LB_TRY(launch_moments_to_order(op, in, so, g0));
      LB_TRY(launch_sweep_pos(op, so, so, g0, true));
    } else {
      LB_TRY(op->w_a2.reserve(n * sizeof(double)));
      LB_TRY(launch_moments_to_discrete(op, in, op->w_a2.as<double>(), g0, 1, 0));
      LB_TRY(launch_sweep_pos(op, op->w_a2.as<const double>(), so, g0));
    }
    return launch_moments_pos(op, so, out);
  }
  LB_TRY(op->w_a2.reserve(n * sizeof(double)));
  double* psi_t = op->w_a2.as<double>();
  LB_TRY(launch_moments_to_discrete(op, in, psi_t, g0, 1, 0));
  LB_TRY(sweep_inplace(op, psi_t, g0));
  if (op->moment_fmt != LB_FMT_FP64) return launch_moments_reduced(op, psi_t, out);
  return launch_moments(op, psi_t, out, 1, op->NENK);
}

// v <- v - D T^-1 M S v: the within-group operator on flux moments of group g0.
int moment_matvec_dev(lb_op* op, double* v, int g0) {
  const int64_t nphi = op->NENK * op->nmom;
  LB_TRY(op->w_phi.reserve((size_t)nphi * sizeof(double)));
  double* tphi = op->w_phi.as<double>();
  LB_TRY(moment_k_dev(op, v, tphi, g0));
  return launch_axpby(op->stream, nphi, 1.0, v, -1.0, tphi, v);
}

// b = D T^-1 q for a discrete source q of group g0.
int moment_rhs_dev(lb_op* op, const double* q, double* b, int g0) {
  const int64_t n = op->NENK * op->NI;
  if (sweep_order_flux(op)) {
    LB_TRY(op->w_so.reserve(n * sizeof(double)));
    LB_TRY(launch_sweep_pos(op, q, op->w_so.as<double>(), g0));
    return launch_moments_pos(op, op->w_so.as<const double>(), b);
  }
  LB_TRY(op->w_c.reserve(n * sizeof(double)));
  double* t = op->w_c.as<double>();
  LB_TRY(launch_copy(op->stream, n, q, t));
  LB_TRY(sweep_inplace(op, t, g0));
  return launch_moments(op, t, b, 1, op->NENK);
}

// Flux-moment GMRES(m) that records every iterate (training data for the
// learned preconditioner).
int gmres_snapshots_dev(lb_op* op, const double* q, const lb_solver_opts* o, double* phi, double* snaps,
                        int max_snaps, int* nsnaps, lb_solve_stats* st) {
  auto t0 = std::chrono::steady_clock::now();
  const int64_t nphi = op->NENK * op->nmom;
  lb_solver_opts opt = o ? *o : lb_solver_opts{LB_SOLVER_GMRES, LB_SPACE_MOMENT, 30, 0, 1e-8, 1e-12, 0};
  const int m = opt.restart > 0 ? opt.restart : 30;
  LB_TRY(op->w_krylov.reserve((size_t)(2 * m + 2) * nphi * sizeof(double)));
  double* V = op->w_krylov.as<double>();
  double* Z = V + (int64_t)(m + 1) * nphi;
  double* rhs = Z + (int64_t)m * nphi;
  LB_TRY(moment_rhs_dev(op, q, rhs, 0));
  LB_TRY(launch_copy(op->stream, nphi, rhs, phi));
  LB_TRY(launch_scale(op->stream, nphi, 1.0 / op->h_sigt[0], phi));
  int count = 0;
  Snapshot snap = [&](const double* sol, const double* Zb, int k, const std::vector<double>& y) -> int {
    if (count >= max_snaps) return LB_OK;
    double* x = snaps + (int64_t)count * nphi;
    LB_TRY(launch_copy(op->stream, nphi, sol, x));
    for (int i = 0; i < k; ++i) LB_TRY(launch_axpby(op->stream, nphi, y[i], Zb + (int64_t)i * nphi, 1.0, x, x));
    ++count;
    return LB_OK;
  };
  Apply A = [&](double* v) { return moment_matvec_dev(op, v, 0); };
  FgmresResult r;
  LB_TRY(fgmres(op, nphi, A, nullptr, rhs, phi, V, Z, nullptr, m, opt.rtol, opt.atol, opt.maxit, opt.verbose, &snap,
                &r));
  LB_HIP(hipStreamSynchronize(op->stream));
  if (nsnaps) *nsnaps = count;
  if (st) {
    *st = lb_solve_stats{};
    st->status = r.status;
    st->iterates = r.iterates;
    st->matvecs = r.matvecs;
    st->sweeps = r.matvecs + 1;
    st->final_norm = r.final_norm;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return r.status == 0 ? LB_OK : r.status;
}

int scatter_dev(lb_op* op, const double* x, double* y, int g0, int G) {
  if (op->scatter_mode == LB_SCATTER_DENSE) {
    if (G != 1 || op->G != 1) {
      set_error("scatter: the dense kernel applies to single-group physics only");
      return LB_E_ARG;
    }
    return launch_scatter_dense(op, x, y, g0, G);
  }
  const int64_t ncols = op->NENK * G;
  LB_TRY(op->w_phi.reserve((size_t)op->nmom * ncols * sizeof(double)));
  LB_TRY(launch_moments(op, x, op->w_phi.as<double>(), G, ncols));
  return launch_moments_to_discrete(op, op->w_phi.as<double>(), y, g0, G, 0);
}

// x <- x - T^-1 K x (BoltzmannOperator.f90:59-104).  K x goes to a separate
// buffer, so x itself serves as the reference's saved copy xw.
int boltzmann_matvec_dev(lb_op* op, double* x, int g0) {
  const int64_t n = op->NENK * op->NI;
  LB_TRY(op->w_c.reserve(n * sizeof(double)));
  double* y = op->w_c.as<double>();
  LB_TRY(scatter_dev(op, x, y, g0, 1));
  LB_TRY(sweep_inplace(op, y, g0));
  return launch_axpby(op->stream, n, 1.0, x, -1.0, y, x);
}

// BoltzmannOp%MatInv (BoltzmannOperator.f90:106-151) for one group g0: s enters
// as the source q and leaves as psi.
int boltzmann_matinv_dev(lb_op* op, double* s, int g0, const lb_solver_opts* o, lb_solve_stats* st) {
  auto t0 = std::chrono::steady_clock::now();
  const int64_t n = op->NENK * op->NI;
  const int64_t nphi = op->NENK * op->nmom;
  lb_solver_opts opt = o ? *o : lb_solver_opts{LB_SOLVER_GMRES, LB_SPACE_PSI, 10, 0, 1e-8, 1e-12, 0};
  if (opt.restart < 1) opt.restart = 10;
  if (opt.kind != LB_SOLVER_GMRES && opt.kind != LB_SOLVER_SI && opt.kind != LB_SOLVER_FGMRES) {
    set_error("BoltzmannOp%MatInv: unknown solver kind");
    return LB_E_ARG;
  }
  lb_solve_stats stats{};
  int sweeps = 0;
  hipStream_t str = op->stream;
  // keep q for the moment-space reconstruction, then b = T^-1 q
  LB_TRY(op->w_b.reserve(n * sizeof(double)));
  LB_TRY(op->w_tq.reserve(n * sizeof(double)));
  double* q = op->w_b.as<double>();
  LB_TRY(launch_copy(str, n, s, q));
  LB_TRY(sweep_inplace(op, s, g0));
  ++sweeps;
  double* tq = op->w_tq.as<double>();   // T^-1 q, kept for the true residual
  LB_TRY(launch_copy(str, n, s, tq));
  const double sig1 = op->h_sigt[(size_t)0 * op->G + g0];  // XS%t(1) of this group

  if (opt.kind == LB_SOLVER_SI) {
    // SourceIteration.f90:34-80 on s = T^-1 q
    LB_TRY(op->w_krylov.reserve(2 * n * sizeof(double)));
    double* w = op->w_krylov.as<double>();
    double* t = w + n;
    LB_TRY(launch_copy(str, n, s, w));
    double ns;
    LB_TRY(dot(op, n, s, s, &ns));
    const double norms = std::sqrt(ns);
    LB_TRY(scatter_dev(op, w, t, g0, 1));
    LB_TRY(sweep_inplace(op, t, g0));
    ++sweeps;
    std::swap(w, t);
    LB_TRY(launch_axpby(str, n, 1.0, w, 1.0, s, s));
    int iter = 1;
    stats.status = LB_E_MAXIT;
    if (opt.verbose) std::printf("             %3d         ----------------------\n", 1);
    while (opt.maxit <= 0 || iter < opt.maxit) {
      ++iter;
      double r2;
      LB_TRY(dot(op, n, w, w, &r2));
      const double res = std::sqrt(r2);
      print_row(opt.verbose, iter, res);
      stats.final_norm = res;
      if (std::isnan(res)) {
        stats.status = LB_E_NAN;
        break;
      }
      if (res <= opt.atol + norms * opt.rtol) {
        stats.status = 0;
        break;
      }
      LB_TRY(scatter_dev(op, w, t, g0, 1));
      LB_TRY(sweep_inplace(op, t, g0));
      ++sweeps;
      std::swap(w, t);
      LB_TRY(launch_axpby(str, n, 1.0, w, 1.0, s, s));
    }
    stats.iterates = iter;
    stats.matvecs = iter - 1;
  } else if (opt.space == LB_SPACE_PSI && opt.kind != LB_SOLVER_FGMRES) {
    const int m = opt.restart;
    LB_TRY(op->w_krylov.reserve((size_t)(m + 2) * n * sizeof(double)));
    double* V = op->w_krylov.as<double>();
    double* sol = V + (int64_t)(m + 1) * n;
    LB_TRY(launch_copy(str, n, s, sol));
    LB_TRY(launch_scale(str, n, 1.0 / sig1, sol));  // xg = s / sigma_t(1)
    GmresResult r;
    Apply A = [&](double* v) {
      ++sweeps;
      return boltzmann_matvec_dev(op, v, g0);
    };
    LB_TRY(gmres(op, n, A, s, sol, V, m, opt.rtol, opt.atol, opt.maxit, opt.verbose, &r));
    LB_TRY(launch_copy(str, n, sol, s));
    stats.status = r.status;
    stats.iterates = r.iterates;
    stats.matvecs = r.matvecs;
    stats.final_norm = r.final_norm;
  } else if (opt.kind == LB_SOLVER_FGMRES) {
    // Flux-moment flexible GMRES with the learned right preconditioner (see the
    // moment-space branch below for the system).
    if (!op->mlp) {
      set_error("FGMRES : no preconditioner set (lb_op_set_preconditioner)");
      return LB_E_NOT_READY;
    }
    const int m = opt.restart;
    LB_TRY(op->w_krylov.reserve((size_t)(2 * m + 4) * nphi * sizeof(double)));
    LB_TRY(op->w_a2.reserve(n * sizeof(double)));
    double* V = op->w_krylov.as<double>();
    double* Z = V + (int64_t)(m + 1) * nphi;
    double* phi = Z + (int64_t)m * nphi;
    double* rhs = phi + nphi;
    double* R = rhs + nphi;
    LB_TRY(launch_moments(op, s, rhs, 1, op->NENK));
    Apply A = [&](double* v) -> int {
      ++sweeps;
      return moment_matvec_dev(op, v, g0);
    };
    Precond Pc = [&](const double* v, double* z) { return mlp_apply(op, v, z, g0); };
    const bool use_net = group_thickness(op, g0) >= op->mlp_min_tau;
    FgmresResult r;
    LB_TRY(fgmres(op, nphi, A, use_net ? &Pc : nullptr, rhs, phi, V, Z, R, m, opt.rtol, opt.atol, opt.maxit,
                  opt.verbose, nullptr, &r, true));
    if (!use_net) r.fallback = -1;
    LB_TRY(launch_copy(str, n, q, s));
    LB_TRY(launch_moments_to_discrete(op, phi, s, g0, 1, 1));
    LB_TRY(sweep_inplace(op, s, g0));
    ++sweeps;
    stats.status = r.status;
    stats.iterates = r.iterates;
    stats.matvecs = r.matvecs;
    stats.final_norm = r.final_norm;
    stats.fallback = r.fallback;
  } else {
    // Flux-moment GMRES: with phi = D psi the within-group system
    // (I - T^-1 M S D) psi = T^-1 q is equivalent to (I - D T^-1 M S) phi = D T^-1 q,
    // and psi = T^-1 (M S phi + q).  Vectors shrink from NI to (L+1)^2 per dof.
    const int m = opt.restart;
    LB_TRY(op->w_krylov.reserve((size_t)(m + 3) * nphi * sizeof(double)));
    LB_TRY(op->w_a2.reserve(n * sizeof(double)));
    double* V = op->w_krylov.as<double>();
    double* phi = V + (int64_t)(m + 1) * nphi;
    double* rhs = phi + nphi;
    LB_TRY(launch_moments(op, s, rhs, 1, op->NENK));
    GmresResult r;
    ApplyK K = [&](const double* in, double* out) -> int {
      ++sweeps;
      return moment_k_dev(op, in, out, g0);
    };
    // x0 = 0 (the psi-space solver keeps the reference's q / sigma_t(1))
    LB_TRY(gmres_shifted(op, nphi, K, rhs, phi, V, m, opt.rtol, opt.atol, opt.maxit, opt.verbose, &r));
    // psi = T^-1 (M S phi + q)
    LB_TRY(launch_copy(str, n, q, s));
    LB_TRY(launch_moments_to_discrete(op, phi, s, g0, 1, 1));
    LB_TRY(sweep_inplace(op, s, g0));
    ++sweeps;
    stats.status = r.status;
    stats.iterates = r.iterates;
    stats.matvecs = r.matvecs;
    stats.final_norm = r.final_norm;
  }
  // true transport residual ||T^-1 q - (I - T^-1 K) psi|| / ||T^-1 q||, with
  // the T^-1 q of the first sweep
  {
    const double* b = tq;
    LB_TRY(launch_copy(str, n, s, q));  // q no longer needed: reuse as A psi
    LB_TRY(boltzmann_matvec_dev(op, q, g0));
    sweeps += 1;
    double bb;
    LB_TRY(dot(op, n, b, b, &bb));
    LB_TRY(launch_axpby(str, n, 1.0, b, -1.0, q, q));
    double rr;
    LB_TRY(dot(op, n, q, q, &rr));
    stats.true_residual = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
  }
  LB_HIP(hipStreamSynchronize(str));
  stats.sweeps = sweeps;
  stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st) *st = stats;
  if (stats.status != 0) {
    const char* what = stats.status == LB_E_MAXIT ? "did not converge in the iteration limit"
                       : stats.status == LB_E_BREAKDOWN ? "anticipated breakdown / division by zero"
                       : "RESIDUAL NORM is NaN";
    set_error(std::string(opt.kind == LB_SOLVER_SI ? "SI : " : "GMRES : ") + what);
    return stats.status;
  }
  return LB_OK;
}

// Multigroup downscatter driver: the group loop of Main.f90:326-476 with the
// energy-coupling source of PIEISource (Main.f90:584-629).  Groups are solved in
// order; each group's flux moments phi_g = D psi_g stay in HBM and feed the
// coupling source of later groups in moment form, replacing the reference's
// per-group angular-flux tapes and its O(G^2) dense K applications.
//   qmom     [G][nmom][NENK] external (first-collision) source moments, or null
//   transfer host [G][G][Nm][L+1]: sigma_l(g' -> g) at [g'][g]; the diagonal
//            g' = g must equal the within-group moments set with the physics
//   phi_out  [G][nmom][NENK] collided flux moments
// Group-batched Gauss-Seidel in energy: batches of B consecutive groups, each
// solved as one coupled flux-moment system
//   (I - D T^-1 M S_B) Phi = D T^-1 q_B,
// S_B the within-group and in-batch downscatter moments, q_B the external source
// plus the downscatter from the groups already solved.  Every operator
// application is one B-group sweep (groups fastest in memory), which costs far
// less per group than B single-group sweeps.  GMRES runs on the system scaled on
// the left by 1 / ||(D T^-1 q_B)_g|| per group, with rtol / sqrt(B), so every
// group's own relative residual meets rtol.  Then psi_B = T^-1 (M S_B Phi + q_B),
// phi_g = D psi_g, and the true transport residual is reported per group.
static int multigroup_batched(lb_op* op, const double* qmom, const double* S, const lb_solver_opts* o,
                              double* phi_out, lb_solve_stats* stats) {
  const int G = op->G;
  const int Bmax = std::min(op->group_batch, G);
  lb_solver_opts opt = o ? *o : lb_solver_opts{LB_SOLVER_GMRES, LB_SPACE_MOMENT, 30, 0, 1e-8, 1e-12, 0};
  if (opt.kind != LB_SOLVER_GMRES) {
    set_error("multigroup solve: group batches use flux-moment GMRES");
    return LB_E_ARG;
  }
  const int m = opt.restart > 0 ? opt.restart : 30;
  const int64_t n1 = op->NENK * op->NI, p1 = op->NENK * op->nmom;
  hipStream_t str = op->stream;
  LB_TRY(op->w_b.reserve((size_t)n1 * Bmax * sizeof(double)));     // q_B
  LB_TRY(op->w_a2.reserve((size_t)n1 * Bmax * sizeof(double)));    // sweep work
  LB_TRY(op->w_mg.reserve((size_t)n1 * Bmax * sizeof(double)));    // psi_B
  LB_TRY(op->w_c.reserve((size_t)n1 * Bmax * sizeof(double)));     // T^-1 q_B
  LB_TRY(op->w_krylov.reserve((size_t)(m + 3) * p1 * Bmax * sizeof(double)));
  LB_TRY(op->w_phi.reserve((size_t)2 * p1 * Bmax * sizeof(double)));
  LB_TRY(op->w_acc.reserve((size_t)p1 * Bmax * sizeof(double)));
  int rc_all = LB_OK;
  for (int g0 = 0; g0 < G; g0 += Bmax) {
    auto t0 = std::chrono::steady_clock::now();
    const int B = std::min(Bmax, G - g0);
    const int64_t n = n1 * B, np = p1 * B;
    double* qB = op->w_b.as<double>();
    double* work = op->w_a2.as<double>();
    double* psi = op->w_mg.as<double>();
    double* tq = op->w_c.as<double>();
    double* V = op->w_krylov.as<double>();
    double* Phi = V + (int64_t)(m + 1) * np;
    double* rhs = Phi + np;
    double* src = op->w_phi.as<double>();
    double* tphi = src + np;
    double* acc = op->w_acc.as<double>();
    int sweeps = 0;
    // q_B = M (qmom + downscatter of the solved groups)
    LB_TRY(launch_batch_source(op, phi_out, S, qmom, g0, B, acc));
    LB_TRY(launch_moments_to_discrete(op, acc, qB, g0, B, 0, false));
    // rhs = W D T^-1 q_B
    LB_TRY(launch_copy(str, n, qB, tq));
    LB_TRY(launch_sweep(op, tq, tq, g0, B));
    ++sweeps;
    LB_TRY(launch_moments(op, tq, rhs, B, op->NENK * B));
    double rn[kMaxGroupBatch], w[kMaxGroupBatch];
    LB_TRY(group_dot(op, np, B, rhs, rhs, rn));
    for (int g = 0; g < B; ++g) w[g] = rn[g] > 0.0 ? 1.0 / std::sqrt(rn[g]) : 1.0;
    LB_TRY(launch_group_scale(op, np, B, w, rhs));
    // x0 = rhs / sigma_t(1) of each group, unscaled
    LB_TRY(launch_copy(str, np, rhs, Phi));
    double x0s[kMaxGroupBatch];
    for (int g = 0; g < B; ++g) x0s[g] = 1.0 / (w[g] * op->h_sigt[(size_t)0 * G + g0 + g]);
    LB_TRY(launch_group_scale(op, np, B, x0s, Phi));
    Apply A = [&](double* v) -> int {
      ++sweeps;
      LB_TRY(launch_batch_transfer(op, S, g0, B, v, src));
      LB_TRY(launch_moments_to_discrete(op, src, work, g0, B, 0, false));
      LB_TRY(launch_sweep(op, work, work, g0, B));
      LB_TRY(launch_moments(op, work, tphi, B, op->NENK * B));
      LB_TRY(launch_axpby(str, np, 1.0, v, -1.0, tphi, v));
      return launch_group_scale(op, np, B, w, v);
    };
    GmresResult r;
    LB_TRY(gmres(op, np, A, rhs, Phi, V, m, opt.rtol / std::sqrt((double)B), opt.atol, opt.maxit, opt.verbose, &r));
    // psi_B = T^-1 (M S_B Phi + q_B), phi = D psi_B
    LB_TRY(launch_batch_transfer(op, S, g0, B, Phi, src));
    LB_TRY(launch_copy(str, n, qB, psi));
    LB_TRY(launch_moments_to_discrete(op, src, psi, g0, B, 1, false));
    LB_TRY(launch_sweep(op, psi, psi, g0, B));
    LB_TRY(launch_moments(op, psi, tphi, B, op->NENK * B));
    LB_TRY(launch_batch_store(op, tphi, g0, B, phi_out));
    // true residual per group: ||T^-1 q - (psi - T^-1 M S_B D psi)|| / ||T^-1 q||
    LB_TRY(launch_batch_transfer(op, S, g0, B, tphi, src));
    LB_TRY(launch_moments_to_discrete(op, src, work, g0, B, 0, false));
    LB_TRY(launch_sweep(op, work, work, g0, B));
    sweeps += 2;
    LB_TRY(launch_axpby(str, n, 1.0, psi, -1.0, work, work));    // (I - T^-1 K) psi
    LB_TRY(launch_axpby(str, n, 1.0, tq, -1.0, work, work));     // residual
    double bb[kMaxGroupBatch], rr[kMaxGroupBatch];
    LB_TRY(group_dot(op, n, B, tq, tq, bb));
    LB_TRY(group_dot(op, n, B, work, work, rr));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int g = 0; g < B; ++g) {
      if (stats) {
        lb_solve_stats& st = stats[g0 + g];
        st = lb_solve_stats{};
        st.status = r.status;
        st.iterates = r.iterates;
        st.matvecs = r.matvecs;
        st.sweeps = g == 0 ? sweeps : 0;   // the batch's sweeps, counted once
        st.final_norm = r.final_norm;
        st.true_residual = bb[g] > 0.0 ? std::sqrt(rr[g] / bb[g]) : std::sqrt(rr[g]);
        st.seconds = g == 0 ? secs : 0.0;
      }
    }
    if (r.status != 0) {
      if (r.status != LB_E_MAXIT && r.status != LB_E_BREAKDOWN && r.status != LB_E_NAN) return r.status;
      rc_all = r.status;
    }
  }
  LB_HIP(hipStreamSynchronize(str));
  return rc_all;
}

// Group-by-group flux-moment GMRES (or flexible GMRES with the learned
// preconditioner) with the residual checks batched over groups.
// Each group g solves (I - D T^-1 M S_g) phi_g = D T^-1 q_g from phi_g = 0, its
// source q_g = M (qmom_g + sum_{g' < g} S_{g'->g} phi_g') built from the groups
// already solved; phi_g is the result.  The true transport residual of
// psi_g = T^-1 (M S_g phi_g + q_g) is
//   T^-1 q_g - (I - T^-1 K_g) psi_g = T^-1 M S_g (D psi_g - phi_g),
// so after all groups are solved, psi and that residual take two sweeps for
// kTailGroups groups at a time instead of two single-group sweeps per group.
static constexpr int kTailGroups = 8;

static int multigroup_moment_gmres(lb_op* op, const double* qmom, const double* S, const lb_solver_opts& opt,
                                   double* phi_out, lb_solve_stats* stats) {
  const int G = op->G;
  const int64_t n = op->NENK * op->NI;
  const int64_t nphi = (int64_t)op->nmom * op->NENK;
  hipStream_t str = op->stream;
  const int m = opt.restart > 0 ? opt.restart : 30;
  const bool flexible = opt.kind == LB_SOLVER_FGMRES;
  if (flexible && !op->mlp) {
    set_error("FGMRES : no preconditioner set (lb_op_set_preconditioner)");
    return LB_E_NOT_READY;
  }
  LB_TRY(op->w_mg.reserve(n * sizeof(double)));
  LB_TRY(op->w_acc.reserve(nphi * sizeof(double)));
  LB_TRY(op->w_a2.reserve(n * sizeof(double)));
  // V (m+1), rhs, and for the flexible solver Z (m) and the carried residual R
  LB_TRY(op->w_krylov.reserve((size_t)(flexible ? 2 * m + 3 : m + 2) * nphi * sizeof(double)));
  const bool so = sweep_order_flux(op);
  if (so) LB_TRY(op->w_so.reserve(n * sizeof(double)));
  double* q = op->w_mg.as<double>();
  double* acc = op->w_acc.as<double>();
  double* V = op->w_krylov.as<double>();
  double* rhs = V + (int64_t)(m + 1) * nphi;
  double* Z = rhs + nphi;
  double* R = Z + (int64_t)m * nphi;
  std::vector<double> tq_norm(G, 0.0);
  std::vector<lb_solve_stats> st(G);
  int rc_all = LB_OK;
  for (int g = 0; g < G; ++g) {
    auto t0 = std::chrono::steady_clock::now();
    LB_TRY(launch_downscatter(op, phi_out, S, g, G, qmom ? qmom + (int64_t)g * nphi : nullptr, acc));
    double* tq = q;                            // T^-1 q_g, in place or in sweep order
    if (so && op->flux_order == LB_FLUX_SWEEP_ORDER_SOURCE) {
      tq = op->w_so.as<double>();
      LB_TRY(launch_moments_to_order(op, acc, tq, g, false));
      LB_TRY(launch_sweep_pos(op, tq, tq, g, true));
    } else if (so) {
      tq = op->w_so.as<double>();
      LB_TRY(launch_moments_to_discrete(op, acc, q, g, 1, 0, false));
      LB_TRY(launch_sweep_pos(op, q, tq, g));
    } else {
      LB_TRY(launch_moments_to_discrete(op, acc, q, g, 1, 0, false));
      LB_TRY(sweep_inplace(op, q, g));
    }
    double tq2;
    LB_TRY(dot(op, n, tq, tq, &tq2));
    tq_norm[g] = std::sqrt(tq2);
    LB_TRY(so ? launch_moments_pos(op, tq, rhs) : launch_moments(op, tq, rhs, 1, op->NENK));
    int sweeps = 1;
    Apply A = [&](double* v) -> int {
      ++sweeps;
      return moment_matvec_dev(op, v, g);
    };
    ApplyK K = [&](const double* in, double* out) -> int {
      ++sweeps;
      return moment_k_dev(op, in, out, g);
    };
    double* phi_g = phi_out + (int64_t)g * nphi;
    // the network only where the training problems showed it pays (lb_op_set_preconditioner_threshold)
    const bool use_net = flexible && group_thickness(op, g) >= op->mlp_min_tau;
    if (use_net) {
      Precond Pc = [&](const double* v, double* z) { return mlp_apply(op, v, z, g); };
      FgmresResult r;
      LB_TRY(fgmres(op, nphi, A, &Pc, rhs, phi_g, V, Z, R, m, opt.rtol, opt.atol, opt.maxit, opt.verbose, nullptr,
                    &r, true));
      st[g].status = r.status;
      st[g].iterates = r.iterates;
      st[g].matvecs = r.matvecs;
      st[g].final_norm = r.final_norm;
      st[g].fallback = r.fallback;
    } else {
      GmresResult r;
      LB_TRY(gmres_shifted(op, nphi, K, rhs, phi_g, V, m, opt.rtol, opt.atol, opt.maxit, opt.verbose, &r));
      st[g].status = r.status;
      st[g].iterates = r.iterates;
      st[g].matvecs = r.matvecs;
      st[g].final_norm = r.final_norm;
      if (flexible) st[g].fallback = -1;   // plain GMRES chosen before the solve
    }
    st[g].sweeps = sweeps;
    st[g].seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st[g].status != 0) rc_all = st[g].status;
  }
  // psi and the true residual, kTailGroups groups per sweep, or fewer when the
  // device cannot hold that many groups' buffers (8 groups at 384k tets x 256
  // directions: 55 GB)
  int B0 = std::min(kTailGroups, G), rc_tail;
  while ((rc_tail = op->w_tail.reserve((size_t)(2 * n + 3 * nphi) * B0 * sizeof(double))) == LB_E_NOMEM && B0 > 1)
    B0 = (B0 + 1) / 2;
  LB_TRY(rc_tail);
  const int64_t nb = n * B0, pb = nphi * B0;
  double* psi_b = op->w_tail.as<double>();
  double* r_b = psi_b + nb;
  double* Phi = r_b + nb;
  double* T = Phi + pb;
  double* tmp = T + pb;
  for (int c0 = 0; c0 < G; c0 += B0) {
    auto t0 = std::chrono::steady_clock::now();
    const int B = std::min(B0, G - c0);
    const int64_t nB = n * B, pB = nphi * B;
    LB_TRY(launch_batch_gather(op, phi_out, c0, B, Phi));
    LB_TRY(launch_batch_source(op, phi_out, S, qmom, c0, B, T));      // qmom + downscatter from g' < c0
    LB_TRY(launch_batch_transfer(op, S, c0, B, Phi, tmp));             // in-chunk g' <= g, within-group included
    LB_TRY(launch_axpby(str, pB, 1.0, tmp, 1.0, T, T));
    LB_TRY(launch_moments_to_discrete(op, T, psi_b, c0, B, 0, false));
    // psi = T^-1 (M S phi + q); the direction-resident kernel at 512 threads, which
    // sweeps 8 groups of 384k tets x 256 directions in 20.8 ms against 22.4 ms at
    // 1024 (graph-replayed level launches: 20.1 ms, but a capture per buffer pair)
    LB_TRY(launch_sweep_direction(op, psi_b, psi_b, c0, B, 512));
    LB_TRY(launch_moments(op, psi_b, tmp, B, op->NENK * B));           // D psi
    LB_TRY(launch_axpby(str, pB, 1.0, tmp, -1.0, Phi, tmp));            // D psi - phi
    LB_TRY(launch_moments_to_discrete(op, tmp, r_b, c0, B, 0, true));  // M S_g (D psi - phi)
    LB_TRY(launch_sweep_direction(op, r_b, r_b, c0, B, 512));
    double rr[kTailGroups];
    LB_TRY(group_dot(op, nB, B, r_b, r_b, rr));
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int b = 0; b < B; ++b) {
      lb_solve_stats& sg = st[c0 + b];
      sg.true_residual = tq_norm[c0 + b] > 0.0 ? std::sqrt(rr[b]) / tq_norm[c0 + b] : std::sqrt(rr[b]);
      sg.sweeps += 2;
      sg.seconds += dt / B;
    }
  }
  op->w_tail.release();
  LB_HIP(hipStreamSynchronize(str));
  if (stats)
    for (int g = 0; g < G; ++g) stats[g] = st[g];
  if (rc_all != LB_OK) {
    set_error(std::string(flexible ? "FGMRES : " : "GMRES : ") +
              (rc_all == LB_E_MAXIT ? "did not converge in the iteration limit"
                                         : rc_all == LB_E_BREAKDOWN ? "anticipated breakdown / division by zero"
                                                                    : "RESIDUAL NORM is NaN"));
  }
  return rc_all;
}

int multigroup_solve_dev(lb_op* op, const double* qmom, const double* transfer, const lb_solver_opts* opts,
                         double* phi_out, lb_solve_stats* stats) {
  const int G = op->G;
  const int64_t n = op->NENK * op->NI;
  const int64_t nphi = (int64_t)op->nmom * op->NENK;
  const size_t nS = (size_t)G * G * op->Nm * (op->L + 1);
  LB_TRY(op->w_S.reserve(nS * sizeof(double)));
  LB_HIP(hipMemcpyAsync(op->w_S.p, transfer, nS * sizeof(double), hipMemcpyHostToDevice, op->stream));
  if (op->group_batch > 1) return multigroup_batched(op, qmom, op->w_S.as<double>(), opts, phi_out, stats);
  if (opts && opts->space == LB_SPACE_MOMENT && (opts->kind == LB_SOLVER_GMRES || opts->kind == LB_SOLVER_FGMRES))
    return multigroup_moment_gmres(op, qmom, op->w_S.as<double>(), *opts, phi_out, stats);
  LB_TRY(op->w_mg.reserve(n * sizeof(double)));
  LB_TRY(op->w_acc.reserve(nphi * sizeof(double)));
  double* psi = op->w_mg.as<double>();
  double* acc = op->w_acc.as<double>();
  int rc_all = LB_OK;
  for (int g = 0; g < G; ++g) {
    LB_TRY(launch_downscatter(op, phi_out, op->w_S.as<double>(), g, G, qmom ? qmom + (int64_t)g * nphi : nullptr,
                              acc));
    LB_TRY(launch_moments_to_discrete(op, acc, psi, g, 1, 0, false));
    lb_solve_stats st{};
    int rc = boltzmann_matinv_dev(op, psi, g, opts, &st);
    if (stats) stats[g] = st;
    if (rc != LB_OK) {
      if (rc != LB_E_MAXIT && rc != LB_E_BREAKDOWN && rc != LB_E_NAN) return rc;
      rc_all = rc;
    }
    LB_TRY(launch_moments(op, psi, phi_out + (int64_t)g * nphi, 1, op->NENK));
  }
  LB_HIP(hipStreamSynchronize(op->stream));
  return rc_all;
}

}  // namespace lb
