// CPU FP64 restatement of the reference transport-solve path (mhyounis/Lionbolt v1.0.0).
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is
// checked against, and the CPU baseline that the benchmark times.  Nothing in
// the product (lionbolt_amd/) may link, load or call it.
//
// Every routine follows one reference routine, cited by file:line, keeping its
// operation order: the quadrature-based mass-eliminated FEM inner products, the
// sorted-face normals, the Kahn-pass sweep order, the sweep-space addressing,
// the encode/decode permutations, the per-visit diagonal rebuild with partial-
// pivot LU, the dense direction x direction scattering kernel, SPARSKIT's
// restarted GMRES with selective re-orthogonalisation, and source iteration.
//
// Pins (what fixes each function without running the reference, which cannot be
// built here: its NittanyPhysics dependency is fetched over the network):
//   fea          closed forms for linear tets (tests/test_oracle_fea.py)
//   normals      invariants sum_f A_f n_f = 0, n_nbr = -n bitwise, unit length
//   sweep order  brute-force longest-path levels + invariants (test_oracle_schedule.py)
//   sweep        dense assembled DFEM solve on small meshes, MMS convergence
//   scattering   addition theorem, P0 closed form, ScatterMMS closed form
//   gmres / si   dense direct solves, textbook Givens, agreement of GMRES and SI
// Parity unpinned: last-bit rounding of gfortran's MATMUL/DOT_PRODUCT/NORM2 and
// of LAPACK dgetri, FMA contraction under -O3 -march=native, which only a build
// of the reference could fix.  This file is compiled with -ffp-contract=off.
//
// Conventions: all arrays are 0-based C order; a boundary neighbour is -1.
// Element e has spatial dofs 4e..4e+3; a space-angle vector is v[sdof + i*NENK].

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <omp.h>

typedef int64_t i64;

namespace {

const double PI = 3.14159265358979323846264338327950288;

// Local faces of a tet, FiniteElementAnalysis.f90:129-132 (0-based).
const int F2K[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

// gfortran's NORM2: scaled sum of squares (the reference BLAS dnrm2 scheme).
double norm2_scaled(const double* x, i64 n) {
  double scale = 0.0, res = 1.0;
  for (i64 i = 0; i < n; ++i) {
    if (x[i] != 0.0) {
      double a = std::fabs(x[i]);
      if (a > scale) {
        double v = scale / a;
        res = 1.0 + res * v * v;
        scale = a;
      } else {
        double v = a / scale;
        res += v * v;
      }
    }
  }
  return scale * std::sqrt(res);
}

void cross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Gauss-Legendre nodes ascending (core::PrepGaussLegendreQuadrature, restated by Newton).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(PI * (i + 0.75) / (n + 0.5));
    double p0 = 1.0, p1 = z, dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      p0 = 1.0;
      p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    p0 = 1.0;
    p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;  // cos-based guesses descend; store ascending
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// Inverse of a small dense matrix by LU with partial pivoting (LAPACK dgetrf + dgetri semantics).
void invert(int n, double* A /*row-major n x n*/) {
  std::vector<double> lu(A, A + n * n), inv(n * n, 0.0);
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(lu[i * n + k]) > std::fabs(lu[p * n + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(lu[k * n + j], lu[p * n + j]);
    for (int i = k + 1; i < n; ++i) {
      lu[i * n + k] /= lu[k * n + k];
      for (int j = k + 1; j < n; ++j) lu[i * n + j] -= lu[i * n + k] * lu[k * n + j];
    }
  }
  for (int c = 0; c < n; ++c) {
    std::vector<double> b(n, 0.0);
    b[c] = 1.0;
    for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) b[i] -= lu[i * n + j] * b[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) b[i] -= lu[i * n + j] * b[j];
      b[i] /= lu[i * n + i];
    }
    for (int i = 0; i < n; ++i) inv[i * n + c] = b[i];
  }
  std::memcpy(A, inv.data(), sizeof(double) * n * n);
}

double tet_sf(int k, const double* r) {  // ShapeFunctions.f90:42-57
  switch (k) {
    case 0: return 1.0 - (r[0] + r[1] + r[2]);
    case 1: return r[0];
    case 2: return r[1];
    default: return r[2];
  }
}
void tet_grad_sf(int k, double* d) {  // ShapeFunctions.f90:59-75
  static const double G[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  d[0] = G[k][0];
  d[1] = G[k][1];
  d[2] = G[k][2];
}
double tri_sf(int k, const double* r) {  // ShapeFunctions.f90:137-150
  switch (k) {
    case 0: return 1.0 - (r[0] + r[1]);
    case 1: return r[0];
    default: return r[1];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Operator state (the union of MeshClass/IP, TransportOp, ScatteringOp fields).
// ---------------------------------------------------------------------------
struct OrcOp {
  i64 NE = 0, NK = 0, NI = 0, NENK = 0;
  int L = 0, Nm = 1;
  std::vector<i64> node, nbr, mat;  // mat 1-based
  std::vector<double> rg;
  // FEM (per element): IC[e][d][k][kp], IFM[e][f][k][kp], IIM[e][f][k][kp]
  std::vector<double> IC, IFM, IIM, nrm /*[e][f][3]*/, area /*[e][f]*/, detJ;
  // angular
  std::vector<double> kdir /*[i][3]*/, w, YY /*[l][ip][i]*/, YYS /*[m][ip][i]*/;
  // schedule (SweepOrder / SweepArrays), 0-based
  std::vector<int32_t> sl /*[i][es]*/, isl /*[i][e]*/, nlev /*[i]*/, lev /*[i][es]*/;
  std::vector<i64> SAoffset /*[i][NE+1]*/, UFoffset /*[i][NE+1]*/;
  std::vector<int8_t> upfaces;
  std::vector<i64> ipsf;
  // physics
  std::vector<double> sigt /*[m]*/, sigs /*[m][l]*/;
  std::vector<std::vector<i64>> mat2sd;
  bool phys = false;
  bool deltadown = false;           // ScatteringOp%BuildDeltaDown
  std::vector<double> xsdd /*[m]*/;
  // K applied in moment form (orc_use_moment_kernel): D [nmom][NI], M [NI][nmom]
  int knmom = 0;
  std::vector<double> kD, kM;
  std::vector<int32_t> kdeg;
  // workspaces
  std::vector<double> xw, sxw;
  // inverse reference mass (FiniteElementAnalysis.f90:167-181), row-major
  double IMr[16] = {0};
};

extern "C" {

// ---------------------------------------------------------------------------
// FiniteElementAnalysis.f90:3-304 with FEAInnerProducts.f90:34-407,
// Jacobians.f90:13-90 and the normal/area part of PostProcess.f90:48-143.
// ---------------------------------------------------------------------------
static void fea(OrcOp* op) {
  const i64 NE = op->NE;
  std::vector<double> xq, wq;
  gauss_legendre(4, xq, wq);  // InitializeFEAQuadrature(4), FiniteElementAnalysis.f90:52
  const int Nq = 4;
  // FEAMassRef tet (FEAInnerProducts.f90:73-98), FEAConvectionRef (:176-200)
  double IMr[16] = {0}, ICr[4][4][3] = {{{0}}};
  for (int ii = 0; ii < Nq; ++ii) {
    double x = 0.5 * xq[ii] + 0.5;
    for (int j = 0; j < Nq; ++j) {
      double y = 0.5 * (1.0 - x) * xq[j] + 0.5 * (1.0 - x);
      for (int kk = 0; kk < Nq; ++kk) {
        double z = 0.5 * (1.0 - x - y) * xq[kk] + 0.5 * (1.0 - x - y);
        double r[3] = {x, y, z};
        double wt = wq[ii] * wq[j] * wq[kk] * 0.125 * (1.0 - x) * (1.0 - x - y);
        for (int kp = 0; kp < 4; ++kp) {
          double ukp = tet_sf(kp, r);
          for (int k = 0; k < 4; ++k) {
            IMr[k * 4 + kp] += wt * tet_sf(k, r) * ukp;
            double d[3];
            tet_grad_sf(k, d);
            for (int c = 0; c < 3; ++c) ICr[k][kp][c] += wt * ukp * d[c];
          }
        }
      }
    }
  }
  invert(4, IMr);  // dgetrf/dgetri, FiniteElementAnalysis.f90:171-180
  std::memcpy(op->IMr, IMr, sizeof(IMr));
  // FEAFaceMassRef(3) (FEAInnerProducts.f90:263-283)
  double IFMr[3][3] = {{0}};
  for (int ii = 0; ii < Nq; ++ii) {
    double x = 0.5 * xq[ii] + 0.5;
    for (int j = 0; j < Nq; ++j) {
      double y = 0.5 * (1.0 - x) * xq[j] + 0.5 * (1.0 - x);
      double r[2] = {x, y};
      for (int kp = 0; kp < 3; ++kp)
        for (int k = 0; k < 3; ++k)
          IFMr[k][kp] += wq[ii] * wq[j] * 0.25 * (1.0 - x) * tri_sf(k, r) * tri_sf(kp, r);
    }
  }

  // Normals and face areas (PostProcess.f90:48-102,135-143): from the sorted global
  // node triple of the face, n = (x3-x1) x (x2-x1), oriented away from the element
  // centroid and normalised; the neighbour receives the exact negation.  Elements are
  // visited in ascending order, so an interior face is computed by its lower-indexed
  // element (the `done` flag of the reference).
  op->nrm.assign(NE * 12, 0.0);
  op->area.assign(NE * 4, 0.0);
  const double* rg = op->rg.data();
  for (i64 e = 0; e < NE; ++e) {
    double com[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) com[c] += rg[op->node[e * 4 + k] * 3 + c];
    for (int c = 0; c < 3; ++c) com[c] /= 4.0;
    for (int f = 0; f < 4; ++f) {
      i64 fn[3] = {op->node[e * 4 + F2K[f][0]], op->node[e * 4 + F2K[f][1]], op->node[e * 4 + F2K[f][2]]};
      std::sort(fn, fn + 3);
      double r1[3], r2[3], n[3];
      for (int c = 0; c < 3; ++c) {
        r1[c] = rg[fn[1] * 3 + c] - rg[fn[0] * 3 + c];
        r2[c] = rg[fn[2] * 3 + c] - rg[fn[0] * 3 + c];
      }
      cross(r2, r1, n);
      op->area[e * 4 + f] = 0.5 * norm2_scaled(n, 3);
      i64 ep = op->nbr[e * 4 + f];
      if (ep >= 0 && ep < e) {  // already done by the neighbour
        int fp = 0;
        while (op->nbr[ep * 4 + fp] != e) ++fp;
        for (int c = 0; c < 3; ++c) op->nrm[e * 12 + f * 3 + c] = -op->nrm[ep * 12 + fp * 3 + c];
        continue;
      }
      double d[3];
      for (int c = 0; c < 3; ++c) d[c] = rg[fn[0] * 3 + c] - com[c];
      double dn = dot3(n, d);
      double normfac = (dn >= 0.0 ? 1.0 : -1.0) / norm2_scaled(n, 3);  // SIGN(ONE, x)
      for (int c = 0; c < 3; ++c) op->nrm[e * 12 + f * 3 + c] = normfac * n[c];
    }
  }

  op->IC.assign(NE * 48, 0.0);
  op->IFM.assign(NE * 64, 0.0);
  op->IIM.assign(NE * 64, 0.0);
  op->detJ.assign(NE, 0.0);
#pragma omp parallel for schedule(static)
  for (i64 e = 0; e < NE; ++e) {
    const i64* nd = &op->node[e * 4];
    // Jacobian (Jacobians.f90:38-90)
    double Jv[3][3], Ji[3][3];
    for (int j = 0; j < 3; ++j)
      for (int c = 0; c < 3; ++c) Jv[c][j] = rg[nd[j + 1] * 3 + c] - rg[nd[0] * 3 + c];
    double detJv = Jv[0][0] * (Jv[1][1] * Jv[2][2] - Jv[1][2] * Jv[2][1]) +
                   Jv[0][1] * (Jv[1][2] * Jv[2][0] - Jv[1][0] * Jv[2][2]) +
                   Jv[0][2] * (Jv[1][0] * Jv[2][1] - Jv[1][1] * Jv[2][0]);
    Ji[0][0] = Jv[1][1] * Jv[2][2] - Jv[1][2] * Jv[2][1];
    Ji[0][1] = Jv[0][2] * Jv[2][1] - Jv[0][1] * Jv[2][2];
    Ji[0][2] = Jv[0][1] * Jv[1][2] - Jv[0][2] * Jv[1][1];
    Ji[1][0] = Jv[1][2] * Jv[2][0] - Jv[1][0] * Jv[2][2];
    Ji[1][1] = Jv[0][0] * Jv[2][2] - Jv[0][2] * Jv[2][0];
    Ji[1][2] = Jv[0][2] * Jv[1][0] - Jv[0][0] * Jv[1][2];
    Ji[2][0] = Jv[1][0] * Jv[2][1] - Jv[1][1] * Jv[2][0];
    Ji[2][1] = Jv[0][1] * Jv[2][0] - Jv[0][0] * Jv[2][1];
    Ji[2][2] = Jv[0][0] * Jv[1][1] - Jv[0][1] * Jv[1][0];
    double JinvT[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) JinvT[a][b] = Ji[b][a] / detJv;  // TRANSPOSE(JvinvT)/detJv
    op->detJ[e] = detJv;
    double IM[4][4];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) IM[a][b] = IMr[a * 4 + b] / detJv;
    // MEFEAConvection (FEAInnerProducts.f90:311-338)
    double tmp[3][4][4];
    for (int kp = 0; kp < 4; ++kp)
      for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) {
          double s = 0.0;
          for (int b = 0; b < 3; ++b) s += JinvT[a][b] * ICr[k][kp][b];
          tmp[a][k][kp] = s;
        }
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 4; ++k)
        for (int kp = 0; kp < 4; ++kp) {
          double s = 0.0;
          for (int j = 0; j < 4; ++j) s += IM[k][j] * tmp[d][j][kp];
          op->IC[e * 48 + d * 16 + k * 4 + kp] = detJv * s;
        }
    for (int f = 0; f < 4; ++f) {
      double normJs = 2.0 * op->area[e * 4 + f];  // FiniteElementAnalysis.f90:278-279
      // MEFEAFaceMass (FEAInnerProducts.f90:340-371)
      double P[4][4] = {{0}};
      for (int kp = 0; kp < 3; ++kp)
        for (int k = 0; k < 3; ++k) P[F2K[f][k]][F2K[f][kp]] = IFMr[k][kp];
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          double s = 0.0;
          for (int j = 0; j < 4; ++j) s += IM[a][j] * P[j][b];
          op->IFM[e * 64 + f * 16 + a * 4 + b] = s * normJs;
        }
      // MEFEAInterfaceMass (FEAInnerProducts.f90:373-407)
      i64 ep = op->nbr[e * 4 + f];
      if (ep < 0) continue;
      double Q[4][4] = {{0}};
      for (int k = 0; k < 3; ++k) {
        int k1 = F2K[f][k];
        for (int kp = 0; kp < 3; ++kp) {
          i64 kgp = nd[F2K[f][kp]];
          int k2 = 0;
          while (op->node[ep * 4 + k2] != kgp) ++k2;
          Q[k1][k2] = IFMr[k][kp];
        }
      }
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          double s = 0.0;
          for (int j = 0; j < 4; ++j) s += IM[a][j] * Q[j][b];
          op->IIM[e * 64 + f * 16 + a * 4 + b] = s * normJs;
        }
    }
  }
}

// ---------------------------------------------------------------------------
// SweepOrder.f90:105-207: per direction, mark interior faces with k.n < -EPSILON
// upstream, then peel levels: each pass adds (in ascending e) every remaining
// element with no remaining dependency; dependencies are cleared after the pass.
// ---------------------------------------------------------------------------
static int sweep_order(OrcOp* op) {
  const i64 NE = op->NE, NI = op->NI;
  const double TOL = 2.220446049250313e-16;  // EPSILON(ONE)
  op->sl.assign(NI * NE, -1);
  op->isl.assign(NI * NE, 0);
  op->lev.assign(NI * NE, 0);
  op->nlev.assign(NI, 0);
  int stagnant = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 i = 0; i < NI; ++i) {
    const double* k = &op->kdir[i * 3];
    std::vector<uint8_t> dep(NE * 4), rem(NE, 1);
    for (i64 e = 0; e < NE; ++e)
      for (int f = 0; f < 4; ++f)
        dep[e * 4 + f] = op->nbr[e * 4 + f] >= 0 && dot3(k, &op->nrm[e * 12 + f * 3]) < -TOL;
    i64 es = 0;
    int level = 0;
    int32_t* sl = &op->sl[i * NE];
    while (true) {
      i64 start = es;
      for (i64 e = 0; e < NE; ++e) {
        if (!rem[e]) continue;
        if (!(dep[e * 4] | dep[e * 4 + 1] | dep[e * 4 + 2] | dep[e * 4 + 3])) {
          sl[es] = (int32_t)e;
          op->isl[i * NE + e] = (int32_t)es;
          op->lev[i * NE + es] = level;
          ++es;
          rem[e] = 0;
        }
      }
      for (i64 p = start; p < es; ++p) {  // clear dependencies after the pass
        i64 e = sl[p];
        for (int f = 0; f < 4; ++f) {
          i64 ep = op->nbr[e * 4 + f];
          if (ep < 0) continue;
          int floc = 0;
          while (op->nbr[ep * 4 + floc] != e) ++floc;  // findloc
          dep[ep * 4 + floc] = 0;
        }
      }
      ++level;
      if (es == NE) break;
      if (es == start) {
#pragma omp atomic write
        stagnant = 1;
        break;
      }
    }
    op->nlev[i] = level;
  }
  return stagnant ? -1 : 0;
}

// SweepArrays.f90:3-94: SAoffset, UFoffset, upfaces, ipsf in sweep space.
static void sweep_arrays(OrcOp* op) {
  const i64 NE = op->NE, NI = op->NI;
  const double TOL = 2.220446049250313e-16;
  op->SAoffset.assign(NI * (NE + 1), 0);
  op->UFoffset.assign(NI * (NE + 1), 0);
  std::vector<i64> cnt(NI, 0);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < NI; ++i) {
    const double* k = &op->kdir[i * 3];
    i64 c = 0;
    for (i64 es = 0; es < NE; ++es) {
      i64 e = op->sl[i * NE + es];
      for (int f = 0; f < 4; ++f)
        if (op->nbr[e * 4 + f] >= 0 && dot3(k, &op->nrm[e * 12 + f * 3]) < -TOL) ++c;
    }
    cnt[i] = c;
  }
  std::vector<i64> base(NI + 1, 0);
  for (i64 i = 0; i < NI; ++i) base[i + 1] = base[i] + cnt[i];
  op->upfaces.assign(base[NI], 0);
  op->ipsf.assign(base[NI], 0);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < NI; ++i) {
    const double* k = &op->kdir[i * 3];
    i64* SA = &op->SAoffset[i * (NE + 1)];
    i64* UF = &op->UFoffset[i * (NE + 1)];
    i64 j = i * op->NENK;  // SAoffset runs over the whole space-angle vector
    for (i64 es = 0; es < NE; ++es) {
      SA[es] = j;
      j += 4;
    }
    SA[NE] = j;
    i64 iupf = base[i];
    for (i64 es = 0; es < NE; ++es) {
      i64 e = op->sl[i * NE + es];
      UF[es] = iupf;
      for (int f = 0; f < 4; ++f) {
        i64 eup = op->nbr[e * 4 + f];
        if (!(eup >= 0 && dot3(k, &op->nrm[e * 12 + f * 3]) < -TOL)) continue;
        op->upfaces[iupf] = (int8_t)f;
        op->ipsf[iupf] = SA[op->isl[i * NE + eup]];
        ++iupf;
      }
    }
    UF[NE] = iupf;
  }
}

// EncodeSweep.f90:3-65 (all elements are 4-node tets, so NK(es) = NK(e)).
static void encode(OrcOp* op, bool enc, double* s) {
  const i64 NE = op->NE, NI = op->NI, NENK = op->NENK;
  double* t = op->xw.data();
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < NI; ++i) {
    i64 ios = i * NENK;
    for (i64 es = 0; es < NE; ++es) {
      i64 e = op->sl[i * NE + es];
      i64 sa = op->SAoffset[i * (NE + 1) + es];
      if (enc)
        for (int k = 0; k < 4; ++k) t[sa + k] = s[ios + 4 * e + k];
      else {
        i64 sa_e = op->SAoffset[i * (NE + 1) + e];
        for (int k = 0; k < 4; ++k) t[sa_e + k] = s[ios + 4 * es + k];
      }
    }
  }
  std::memcpy(s, t, sizeof(double) * NENK * NI);
}

// LocalLU (TransportAlgebra.f90:173-246): Doolittle with partial pivoting.
static void local_lu(double A[4][4], double b[4]) {
  const int n = 4;
  for (int k = 0; k < n - 1; ++k) {
    double maxval = std::fabs(A[k][k]);
    int pivot = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i][k]) > maxval) {
        maxval = std::fabs(A[i][k]);
        pivot = i;
      }
    if (pivot != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k][j], A[pivot][j]);
      std::swap(b[k], b[pivot]);
    }
    for (int i = k + 1; i < n; ++i) {
      double factor = A[i][k] / A[k][k];
      A[i][k] = factor;
      for (int j = k + 1; j < n; ++j) A[i][j] = A[i][j] - factor * A[k][j];
      b[i] = b[i] - factor * b[k];
    }
  }
  double x[4];
  x[n - 1] = b[n - 1] / A[n - 1][n - 1];
  for (int i = n - 2; i >= 0; --i) {
    x[i] = b[i];
    for (int j = i + 1; j < n; ++j) x[i] = x[i] - A[i][j] * x[j];
    x[i] = x[i] / A[i][i];
  }
  for (int i = 0; i < n; ++i) b[i] = x[i];
}

// Sweep (TransportAlgebra.f90:29-106) with ApplyInvertedDiagonal (:108-171).
static void sweep(OrcOp* op, double* s) {
  const i64 NE = op->NE, NI = op->NI;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < NI; ++i) {
    const double* lk = &op->kdir[i * 3];
    for (i64 es = 0; es < NE; ++es) {
      i64 e = op->sl[i * NE + es];
      double attn = op->sigt[op->mat[e] - 1];
      i64 ips = op->SAoffset[i * (NE + 1) + es];
      i64 u0 = op->UFoffset[i * (NE + 1) + es], u1 = op->UFoffset[i * (NE + 1) + es + 1];
      for (i64 u = u0; u < u1; ++u) {
        int f = op->upfaces[u];
        i64 ipsf = op->ipsf[u];
        double kdotn = dot3(lk, &op->nrm[e * 12 + f * 3]);
        const double* IIM = &op->IIM[e * 64 + f * 16];
        double tmp[4];
        for (int a = 0; a < 4; ++a) {  // -kdotn * MATMUL(IIM, s_up)
          double m = 0.0;
          for (int b = 0; b < 4; ++b) m += IIM[a * 4 + b] * s[ipsf + b];
          tmp[a] = -kdotn * m;
        }
        for (int a = 0; a < 4; ++a) s[ips + a] = s[ips + a] + tmp[a];
      }
      double T[4][4] = {{0}};
      for (int k = 0; k < 4; ++k) T[k][k] = attn;  // M term
      for (int d = 0; d < 3; ++d)                  // G term
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) T[a][b] = T[a][b] - lk[d] * op->IC[e * 48 + d * 16 + a * 4 + b];
      for (int f = 0; f < 4; ++f) {  // outflow face term
        double kdotn = dot3(lk, &op->nrm[e * 12 + f * 3]);
        if (kdotn < 0.0) continue;
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) T[a][b] = T[a][b] + kdotn * op->IFM[e * 64 + f * 16 + a * 4 + b];
      }
      double b[4] = {s[ips], s[ips + 1], s[ips + 2], s[ips + 3]};
      local_lu(T, b);
      for (int a = 0; a < 4; ++a) s[ips + a] = b[a];
    }
  }
}

// SNPrep.f90:106-133: YY(ip,i,l) = w_ip (2l+1)/(4 pi) P_l(k_i.k_ip), diagonal P = 1.
static void build_yy(OrcOp* op) {
  const i64 NI = op->NI;
  const int L = op->L;
  op->YY.assign((L + 1) * NI * NI, 0.0);
  for (i64 ip = 0; ip < NI; ++ip) {
    for (i64 i = 0; i <= ip; ++i) {
      double x = (i == ip) ? 1.0 : dot3(&op->kdir[i * 3], &op->kdir[ip * 3]);
      double p0 = 1.0, p1 = x;
      for (int l = 0; l <= L; ++l) {
        double P;
        if (l == 0) P = 1.0;
        else if (l == 1) P = x;
        else {
          double p2 = ((2 * l - 1) * x * p1 - (l - 1) * p0) / l;
          p0 = p1;
          p1 = p2;
          P = p2;
        }
        if (i == ip) P = 1.0;
        double v = (2 * l + 1) * P / (4.0 * PI);
        op->YY[(l * NI + ip) * NI + i] = v;
        op->YY[(l * NI + i) * NI + ip] = v;
      }
    }
  }
  for (int l = 0; l <= L; ++l)
    for (i64 ip = 0; ip < NI; ++ip)
      for (i64 i = 0; i < NI; ++i) op->YY[(l * NI + ip) * NI + i] *= op->w[ip];
}

void* orc_op_create(i64 NE, i64 NK, const double* rg, const int64_t* node, const int64_t* nbr,
                    const int64_t* mat, i64 NI, const double* kdir, const double* w, int L) {
  OrcOp* op = new OrcOp;
  op->NE = NE;
  op->NK = NK;
  op->NI = NI;
  op->NENK = 4 * NE;
  op->L = L;
  op->rg.assign(rg, rg + NK * 3);
  op->node.assign(node, node + NE * 4);
  op->nbr.assign(nbr, nbr + NE * 4);
  op->mat.assign(mat, mat + NE);
  op->Nm = 1;
  for (i64 e = 0; e < NE; ++e) op->Nm = std::max<int>(op->Nm, (int)mat[e]);
  op->kdir.assign(kdir, kdir + NI * 3);
  op->w.assign(w, w + NI);
  fea(op);
  if (sweep_order(op) != 0) {
    delete op;
    return nullptr;
  }
  sweep_arrays(op);
  build_yy(op);
  // mat2sd (PostProcess.f90:433-453)
  op->mat2sd.assign(op->Nm, {});
  for (i64 e = 0; e < NE; ++e)
    for (int k = 0; k < 4; ++k) op->mat2sd[op->mat[e] - 1].push_back(4 * e + k);
  op->xw.assign(op->NENK * NI, 0.0);
  return op;
}

void orc_op_destroy(void* h) { delete static_cast<OrcOp*>(h); }

// ScatteringOp%Physics (ScatteringOperator.f90:215-269) and TransportOp%Physics.
void orc_set_physics(void* h, const double* sigt /*[Nm]*/, const double* sigs /*[Nm][L+1]*/) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 NI = op->NI;
  op->sigt.assign(sigt, sigt + op->Nm);
  op->sigs.assign(sigs, sigs + op->Nm * (op->L + 1));
  op->YYS.assign(op->Nm * NI * NI, 0.0);
  for (int m = 0; m < op->Nm; ++m)
    for (int l = 0; l <= op->L; ++l)
      for (i64 j = 0; j < NI * NI; ++j) op->YYS[m * NI * NI + j] += op->YY[l * NI * NI + j] * op->sigs[m * (op->L + 1) + l];
  op->phys = true;
  op->deltadown = false;
}

void orc_get_geometry(void* h, double* IC, double* IFM, double* IIM, double* nrm, double* area, double* detJ) {
  OrcOp* op = static_cast<OrcOp*>(h);
  if (IC) std::memcpy(IC, op->IC.data(), sizeof(double) * op->IC.size());
  if (IFM) std::memcpy(IFM, op->IFM.data(), sizeof(double) * op->IFM.size());
  if (IIM) std::memcpy(IIM, op->IIM.data(), sizeof(double) * op->IIM.size());
  if (nrm) std::memcpy(nrm, op->nrm.data(), sizeof(double) * op->nrm.size());
  if (area) std::memcpy(area, op->area.data(), sizeof(double) * op->area.size());
  if (detJ) std::memcpy(detJ, op->detJ.data(), sizeof(double) * op->detJ.size());
}

void orc_get_schedule(void* h, int32_t* sl, int32_t* nlev, int32_t* lev) {
  OrcOp* op = static_cast<OrcOp*>(h);
  if (sl) std::memcpy(sl, op->sl.data(), sizeof(int32_t) * op->sl.size());
  if (nlev) std::memcpy(nlev, op->nlev.data(), sizeof(int32_t) * op->nlev.size());
  if (lev) std::memcpy(lev, op->lev.data(), sizeof(int32_t) * op->lev.size());
}

void orc_get_yy(void* h, double* yy) {
  OrcOp* op = static_cast<OrcOp*>(h);
  std::memcpy(yy, op->YY.data(), sizeof(double) * op->YY.size());
}

// TransportOp%MatInv 'SWEEP' (TransportOperator.f90:101-147): encode, sweep, decode.
void orc_transport_matinv(void* h, double* s) {
  OrcOp* op = static_cast<OrcOp*>(h);
  encode(op, true, s);
  sweep(op, s);
  encode(op, false, s);
}

// Sweep without the encode/decode permutations, on a vector already in sweep
// order (for timing the sweep alone).
void orc_sweep_only(void* h, double* s) { sweep(static_cast<OrcOp*>(h), s); }

// ScatteringOp%BuildDeltaDown (ScatteringOperator.f90:291-319): the next
// scattering products apply XSdd(mat) times the identity in angle, until
// orc_set_physics builds the Legendre kernel again (ScatteringOperator.f90:231).
void orc_build_delta_down(void* h, const double* xsdd /*[Nm]*/) {
  OrcOp* op = static_cast<OrcOp*>(h);
  op->xsdd.assign(xsdd, xsdd + op->Nm);
  op->deltadown = true;
}

// SNMGXS (ScatteringAlgebra.f90:13-55): x(NENK,NI) <- x . YYS; dgemm for one
// material, per-dof dgemv('T') over mat2sd otherwise.  With a delta-down kernel,
// ApplyDeltaDown (ScatteringAlgebra.f90:62-90): x(sd + i NENK) *= XSdd(mat).
void orc_scatter_moment_matvec(void* h, const double* D, const double* M, const int32_t* deg, int nmom, double* x);

void orc_scatter_matvec(void* h, double* x) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 NI = op->NI, NENK = op->NENK;
  if (op->deltadown) {
    for (int m = 0; m < op->Nm; ++m)
      for (i64 sd : op->mat2sd[m])
        for (i64 i = 0; i < NI; ++i) x[sd + i * NENK] *= op->xsdd[m];
    return;
  }
  if (op->knmom > 0) {
    orc_scatter_moment_matvec(h, op->kD.data(), op->kM.data(), op->kdeg.data(), op->knmom, x);
    return;
  }
  double* xw = op->xw.data();
  std::vector<i64> sd2m(NENK);
  for (int m = 0; m < op->Nm; ++m)
    for (i64 sd : op->mat2sd[m]) sd2m[sd] = m;
#pragma omp parallel
  {
    std::vector<double> xt(NI), yt(NI);
#pragma omp for schedule(static)
    for (i64 sd = 0; sd < NENK; ++sd) {
      const double* Y = &op->YYS[sd2m[sd] * NI * NI];
      for (i64 ip = 0; ip < NI; ++ip) xt[ip] = x[sd + ip * NENK];
      for (i64 i = 0; i < NI; ++i) {
        double acc = 0.0;
        for (i64 ip = 0; ip < NI; ++ip) acc += xt[ip] * Y[ip * NI + i];
        yt[i] = acc;
      }
      for (i64 i = 0; i < NI; ++i) xw[sd + i * NENK] = yt[i];
    }
  }
  std::memcpy(x, xw, sizeof(double) * NENK * NI);  // xwp = xw
}

// K x in moment form, for the CPU baseline only (no reference routine: the
// reference always applies the dense kernel above).  phi[lm][sd] =
// sum_i D[lm][i] x[i][sd]; x[i][sd] <- sum_lm M[i][lm] sigma_{deg[lm]}(mat(sd))
// phi[lm][sd].  With D = w Y and M = Y^T over real orthonormal harmonics this is
// the dense kernel by the addition theorem; timing it keeps the O(NI^2) ->
// O(NI nmom) change out of the GPU/CPU ratio.
void orc_scatter_moment_matvec(void* h, const double* D /*[nmom][NI]*/, const double* M /*[NI][nmom]*/,
                               const int32_t* deg, int nmom, double* x) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 NI = op->NI, NENK = op->NENK;
  const int L1 = op->L + 1;
  const i64 CH = 64;  // spatial dofs per block
#pragma omp parallel
  {
    std::vector<double> phi((size_t)nmom * CH), y((size_t)NI * CH);
#pragma omp for schedule(static)
    for (i64 s0 = 0; s0 < NENK; s0 += CH) {
      const i64 nc = std::min(CH, NENK - s0);
      std::fill(phi.begin(), phi.end(), 0.0);
      for (i64 i = 0; i < NI; ++i) {
        const double* xi = x + i * NENK + s0;
        for (int lm = 0; lm < nmom; ++lm) {
          const double d = D[lm * NI + i];
          double* p = &phi[(size_t)lm * CH];
          for (i64 c = 0; c < nc; ++c) p[c] += d * xi[c];
        }
      }
      for (int lm = 0; lm < nmom; ++lm)
        for (i64 c = 0; c < nc; ++c)
          phi[(size_t)lm * CH + c] *= op->sigs[(op->mat[(s0 + c) >> 2] - 1) * L1 + deg[lm]];
      for (i64 i = 0; i < NI; ++i) {
        double* yi = &y[(size_t)i * CH];
        std::fill(yi, yi + CH, 0.0);
        for (int lm = 0; lm < nmom; ++lm) {
          const double mm = M[i * nmom + lm];
          const double* p = &phi[(size_t)lm * CH];
          for (i64 c = 0; c < nc; ++c) yi[c] += mm * p[c];
        }
      }
      for (i64 i = 0; i < NI; ++i) std::memcpy(x + i * NENK + s0, &y[(size_t)i * CH], sizeof(double) * nc);
    }
  }
}

// Makes every later K application (scatter matvec, Boltzmann matvec, GMRES, SI)
// use the moment form above with these D, M and degrees; nmom = 0 returns to
// the dense kernel.  Not a reference routine: at 256 directions the dense kernel
// costs 2 NI^2 flops per dof, which makes full-size oracle solves take many
// minutes; the two forms agree to rounding by the addition theorem (pinned in
// tests/test_oracle_scatter_krylov.py).
void orc_use_moment_kernel(void* h, const double* D /*[nmom][NI]*/, const double* M /*[NI][nmom]*/,
                           const int32_t* deg, int nmom) {
  OrcOp* op = static_cast<OrcOp*>(h);
  op->knmom = nmom > 0 ? nmom : 0;
  if (op->knmom == 0) return;
  op->kD.assign(D, D + (size_t)nmom * op->NI);
  op->kM.assign(M, M + (size_t)nmom * op->NI);
  op->kdeg.assign(deg, deg + nmom);
}

// BoltzmannOp%MatVec (BoltzmannOperator.f90:59-104), Tprecond: x <- x - T^{-1} K x.
void orc_boltzmann_matvec(void* h, double* x) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 n = op->NENK * op->NI;
  op->sxw.assign(x, x + n);
  orc_scatter_matvec(h, x);
  orc_transport_matinv(h, x);
#pragma omp parallel for schedule(static)
  for (i64 j = 0; j < n; ++j) x[j] = op->sxw[j] - x[j];
}

// distdot (distdot.f:1-7, a BLAS ddot): products summed in index order inside
// fixed blocks of 32,768 entries, then the block sums in block order, so the
// result does not depend on the thread count (one block: the plain loop).
static double ddot(i64 n, const double* a, const double* b) {
  const i64 B = i64(1) << 15;
  const i64 nb = (n + B - 1) / B;
  if (nb <= 1) {
    double s = 0.0;
    for (i64 i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  }
  std::vector<double> part(nb);
#pragma omp parallel for schedule(static)
  for (i64 j = 0; j < nb; ++j) {
    const i64 i1 = std::min(n, (j + 1) * B);
    double s = 0.0;
    for (i64 i = j * B; i < i1; ++i) s += a[i] * b[i];
    part[j] = s;
  }
  double s = 0.0;
  for (i64 j = 0; j < nb; ++j) s += part[j];
  return s;
}

// y = x (elementwise, threaded)
static void vcopy(i64 n, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (i64 j = 0; j < n; ++j) y[j] = x[j];
}

// SPARSKIT gmres (iters.f:1901-2206) with mgsro (:3465-3586) and givens (:3225-3260),
// driven as GMRESWrapper does (SPARSKIT.f90:20-181): ipar(3)=2, fpar(1:2)=rtol,atol,
// no preconditioner, no iteration limit unless maxit > 0.  `hist` receives fpar(5)
// at every reverse-communication return; the return value is the wrapper's ipar(1)
// at exit (0 converged, -1 iteration limit, -3 breakdown); *iters is the wrapper's
// "Total number of iterates" and *nmatvec the operator applications.
int orc_gmres(void* h, int m, const double* rhs, double* sol, double rtol, double atol, int maxit,
              double* hist, int hist_cap, int* iters, int* nmatvec) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 n = op->NENK * op->NI;
  std::vector<double> V((i64)(m + 1) * n, 0.0);
  std::vector<double> H((m + 1) * (m + 2) / 2 + 2, 0.0), c(m + 1), s(m + 1), rs(m + 2);
  // Packed Hessenberg: column k (1-based) holds rows 1..k+1 at H[k(k-1)/2 + (row-1)].
  auto hcol = [&](int k) { return &H[(i64)k * (k - 1) / 2]; };
  int iter = 0, nmv = 0;
  double fpar3 = 0.0, fpar4 = 0.0, fpar5 = 0.0;
  int status = 1;
  auto record = [&](double v) {
    if (iter < hist_cap) hist[iter] = v;
    ++iter;
  };
  record(0.0);  // first return: request A*sol, fpar(5) cleared by bisinit
  bool first = true;
  while (true) {
    // restart (label 100/10/20): r = rhs - A sol
    double* v0 = V.data();
    vcopy(n, sol, v0);
    orc_boltzmann_matvec(h, v0);
    ++nmv;
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < n; ++j) v0[j] = rhs[j] - v0[j];
    double alpha = std::sqrt(ddot(n, v0, v0));
    if (first) {
      fpar4 = rtol * std::sqrt(ddot(n, rhs, rhs)) + atol;
      fpar3 = alpha;
      first = false;
    }
    fpar5 = alpha;
    rs[1] = alpha;
    if (alpha <= fpar4) {
      status = 0;
      record(fpar5);
      break;
    }
    alpha = 1.0 / alpha;
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < n; ++j) v0[j] = alpha * v0[j];
    int k = 0;
    int ierr = 0;
    while (true) {
      ++k;
      record(fpar5);  // return requesting A * v_{k-1}
      double* vk = &V[(i64)k * n];
      vcopy(n, &V[(i64)(k - 1) * n], vk);
      orc_boltzmann_matvec(h, vk);
      ++nmv;
      // mgsro(.false., n, n, k+1, k+1, ...): orthogonalise v_k against v_0..v_{k-1}
      double* hh = hcol(k);
      double nrm0 = ddot(n, vk, vk);
      double thr = nrm0 * 0.98;
      if (!(nrm0 > 0.0)) {
        ierr = nrm0 <= 0.0 ? -1 : -2;
      } else {
        for (int i = 0; i < k; ++i) {
          const double* vi = &V[(i64)i * n];
          double fct = ddot(n, vk, vi);
          hh[i] = fct;
#pragma omp parallel for schedule(static)
          for (i64 j = 0; j < n; ++j) vk[j] = vk[j] - fct * vi[j];
          if (fct * fct > thr) {
            fct = ddot(n, vk, vi);
            hh[i] = hh[i] + fct;
#pragma omp parallel for schedule(static)
            for (i64 j = 0; j < n; ++j) vk[j] = vk[j] - fct * vi[j];
          }
          nrm0 = nrm0 - hh[i] * hh[i];
          if (nrm0 < 0.0) nrm0 = 0.0;
          thr = nrm0 * 0.98;
        }
        double nrm1 = std::sqrt(ddot(n, vk, vk));
        hh[k] = nrm1;
        if (nrm1 <= 0.0) {
          ierr = -3;
        } else {
          double fct = 1.0 / nrm1;
#pragma omp parallel for schedule(static)
          for (i64 j = 0; j < n; ++j) vk[j] = vk[j] * fct;
        }
      }
      if (ierr < 0) break;
      // apply previous rotations, then a new one (iters.f:2069-2085)
      for (int i = 1; i <= k - 1; ++i) {
        double a = hh[i - 1];
        double ci = c[i], si = s[i];
        hh[i - 1] = ci * a + si * hh[i];
        hh[i] = ci * hh[i] - si * a;
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
      s[k] = ss;
      double a2 = -ss * rs[k];
      rs[k] = cc * rs[k];
      rs[k + 1] = a2;
      fpar5 = std::fabs(a2);
      if (!(k < m && !(fpar5 <= fpar4) && (maxit <= 0 || nmv < maxit))) break;
    }
    // label 200: upper-triangular solve and update (iters.f:2100-2155)
    int kk = k;
    bool singular = false;
    while (kk > 0 && hcol(kk)[kk - 1] == 0.0) --kk;
    if (kk == 0) singular = true;
    if (!singular) {
      std::vector<double> y(rs.begin(), rs.end());
      y[kk] = y[kk] / hcol(kk)[kk - 1];
      for (int i = kk - 1; i >= 1; --i) {
        const double* col = hcol(i + 1);
        for (int ii = 1; ii <= i; ++ii) y[ii] = y[ii] - y[i + 1] * col[ii - 1];
        y[i] = y[i] / hcol(i)[i - 1];
      }
      double* u = V.data();
      const double y1 = y[1];
#pragma omp parallel for schedule(static)
      for (i64 j = 0; j < n; ++j) u[j] = u[j] * y1;
      for (int i = 1; i <= kk - 1; ++i) {
        const double* vi = &V[(i64)i * n];
        double yi = y[i + 1];
#pragma omp parallel for schedule(static)
        for (i64 j = 0; j < n; ++j) u[j] = u[j] + yi * vi[j];
      }
#pragma omp parallel for schedule(static)
      for (i64 j = 0; j < n; ++j) sol[j] = sol[j] + u[j];
    }
    if (ierr != 0 || singular) {
      status = -3;
      record(fpar5);
      break;
    }
    // label 70: restart or terminate (fpar(6) = fpar(5) since ipar(3) = 2)
    if ((maxit <= 0 || nmv < maxit) && fpar5 > fpar4) {
      record(fpar5);  // return requesting A*sol for the restart
      continue;
    }
    status = (fpar5 <= fpar4) ? 0 : -1;
    record(fpar5);
    break;
  }
  (void)fpar3;
  if (iters) *iters = iter;
  if (nmatvec) *nmatvec = nmv;
  return status;
}

// SourceIteration (SourceIteration.f90:3-82) on s = T^{-1} q.
int orc_source_iteration(void* h, double* s, double rtol, double atol, int maxit, double* hist,
                         int hist_cap, int* iters) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 n = op->NENK * op->NI;
  std::vector<double> w(s, s + n);
  double norms = norm2_scaled(s, n);
  orc_scatter_matvec(h, w.data());
  orc_transport_matinv(h, w.data());
#pragma omp parallel for schedule(static)
  for (i64 j = 0; j < n; ++j) s[j] = s[j] + w[j];
  int iter = 1;
  int status = -1;
  while (maxit <= 0 || iter < maxit) {
    ++iter;
    double res = norm2_scaled(w.data(), n);
    if (iter - 2 < hist_cap) hist[iter - 2] = res;
    if (res <= atol + norms * rtol) {
      status = 0;
      break;
    }
    orc_scatter_matvec(h, w.data());
    orc_transport_matinv(h, w.data());
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < n; ++j) s[j] = s[j] + w[j];
  }
  if (iters) *iters = iter;
  return status;
}

// BoltzmannOp%MatInv (BoltzmannOperator.f90:106-151): s <- T^{-1} s, then GMRES
// from x0 = s/sigma_t(material 1) or source iteration.  kind 0 = GMRES, 1 = SI.
int orc_boltzmann_matinv(void* h, double* s, int kind, int m, double rtol, double atol, int maxit,
                         double* hist, int hist_cap, int* iters, int* nmatvec) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 n = op->NENK * op->NI;
  orc_transport_matinv(h, s);
  if (kind == 0) {
    std::vector<double> xg(n);
    const double st0 = op->sigt[0];
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < n; ++j) xg[j] = s[j] / st0;
    int st = orc_gmres(h, m, s, xg.data(), rtol, atol, maxit, hist, hist_cap, iters, nmatvec);
    std::memcpy(s, xg.data(), sizeof(double) * n);
    return st;
  }
  if (nmatvec) *nmatvec = 0;
  return orc_source_iteration(h, s, rtol, atol, maxit, hist, hist_cap, iters);
}

// ---------------------------------------------------------------------------
// Planar external beam: the uncollided fluence at tetrahedral quadrature points
// (init_ExternalBeam.f90:62-95, InField.f90, RayTracing.f90:24-205,
// PrepFluence.f90:42-58, FluenceQuadrature.f90:3-33), its projection on the
// nodal basis and the first-collision source (FCSSource.f90:134-215), and the
// uncollided angular fluence (AddAngularFluence.f90:109-166).
// ---------------------------------------------------------------------------

// core::PrepTetrahedralQuadrature (NittanyPhysics, not vendored): restated as
// the symmetric 4-point rule exact for quadratics, barycentric points
// (a, b, b, b) and their permutations with a = (5 + 3 sqrt 5) / 20 and
// b = (5 - sqrt 5) / 20, weight 1/24 each on the reference tetrahedron of volume
// 1/6; point kq is the one nearest node kq.  Its points are parity unpinned.
static void tet_quadrature(double cq[4][3], double wq[4]) {
  const double a = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0, b = (5.0 - std::sqrt(5.0)) / 20.0;
  for (int kq = 0; kq < 4; ++kq) {
    for (int c = 0; c < 3; ++c) cq[kq][c] = (kq == c + 1) ? a : b;   // (xi, eta, zeta) = weights of nodes 2..4
    wq[kq] = 1.0 / 24.0;
  }
}

struct OrcBeam {
  double axis[3], origin[3], e1[3], e2[3];
  int cutout;        // 0 none, 1 circle, 2 rectangle
  double params[4];  // FieldGeometry%cutoutparams
  double n;          // particles in the beam
};

// InField.f90:3-62, planar beams.
static bool in_field(const OrcBeam& B, const double* r) {
  const double TOL = 1.0e-12;
  if (B.cutout == 0) return true;
  double u[3], v[3];
  for (int c = 0; c < 3; ++c) u[c] = r[c] - B.origin[c];
  const double ua = dot3(u, B.axis);
  for (int c = 0; c < 3; ++c) v[c] = u[c] - ua * B.axis[c];
  if (B.cutout == 1) return norm2_scaled(v, 3) <= B.params[1] + TOL;
  if (!(std::fabs(dot3(v, B.e1)) <= 0.5 * B.params[2] + TOL)) return false;
  return std::fabs(dot3(v, B.e2)) <= 0.5 * B.params[3] + TOL;
}

// A region for ray tracing: a connected set of elements of one material
// (ReadGMSH.f90:259-312 joins mesh blocks of equal material that touch) and the
// faces enclosing it, each as (the first element sharing it, its local face)
// (PostProcess.f90:106-130; one material: every boundary face, ReadGMSH.f90:261-265).
struct OrcRegion {
  int mat;
  std::vector<std::pair<i64, int>> faces;
};

static void build_regions(OrcOp* op, std::vector<OrcRegion>& reg) {
  const i64 NE = op->NE;
  reg.clear();
  if (op->Nm == 1) {
    reg.push_back({1, {}});
    for (i64 e = 0; e < NE; ++e)
      for (int f = 0; f < 4; ++f)
        if (op->nbr[e * 4 + f] < 0) reg[0].faces.push_back({e, f});
    return;
  }
  std::vector<i64> rid(NE, -1);
  for (i64 s = 0; s < NE; ++s) {
    if (rid[s] >= 0) continue;
    const i64 r = (i64)reg.size();
    reg.push_back({(int)op->mat[s], {}});
    std::vector<i64> stack{s};
    rid[s] = r;
    while (!stack.empty()) {
      const i64 e = stack.back();
      stack.pop_back();
      for (int f = 0; f < 4; ++f) {
        const i64 ep = op->nbr[e * 4 + f];
        if (ep >= 0 && rid[ep] < 0 && op->mat[ep] == op->mat[e]) {
          rid[ep] = r;
          stack.push_back(ep);
        }
      }
    }
  }
  for (i64 e = 0; e < NE; ++e)
    for (int f = 0; f < 4; ++f) {
      const i64 ep = op->nbr[e * 4 + f];
      if (ep < 0) {
        reg[rid[e]].faces.push_back({e, f});
      } else if (ep > e && rid[ep] != rid[e]) {
        reg[rid[e]].faces.push_back({e, f});
        reg[rid[ep]].faces.push_back({e, f});
      }
    }
}

// TraceVector (RayTracing.f90:207-369): path length in each material of the
// segment from R0 to R0 + r, from its intersections (Moller-Trumbore with the
// reference's tolerances) with the faces enclosing each region.  Returns false
// when a one-material trace finds no entry face (the reference stops there).
static bool trace_vector(const OrcOp* op, const std::vector<OrcRegion>& reg, const double* R0, const double* r,
                         double* pathlen) {
  const double TOL = 1.0e-8;
  const int Nm = op->Nm;
  for (int m = 0; m < Nm; ++m) pathlen[m] = 0.0;
  const double length = norm2_scaled(r, 3);
  bool success = false;
  double t = 0.0;
  std::vector<double> llist;
  for (const OrcRegion& R : reg) {
    const int mat = Nm > 1 ? R.mat : 1;
    llist.clear();
    for (const auto& ef : R.faces) {
      const i64 e = ef.first;
      const int f = ef.second;
      i64 fn[3] = {op->node[e * 4 + F2K[f][0]], op->node[e * 4 + F2K[f][1]], op->node[e * 4 + F2K[f][2]]};
      std::sort(fn, fn + 3);
      const double* v0 = &op->rg[fn[0] * 3];
      const double* v1 = &op->rg[fn[1] * 3];
      const double* v2 = &op->rg[fn[2] * 3];
      double edge1[3], edge2[3], nref[3], pvec[3], tvec[3], qvec[3];
      for (int c = 0; c < 3; ++c) {
        edge1[c] = v1[c] - v0[c];
        edge2[c] = v2[c] - v0[c];
      }
      const double* n = &op->nrm[e * 12 + f * 3];
      double det = dot3(n, r);
      if (Nm == 1) {
        if (det > -TOL) continue;
      } else if (std::fabs(det) < TOL) {
        continue;
      }
      cross(edge1, edge2, nref);
      if (dot3(n, nref) < 0.0)
        for (int c = 0; c < 3; ++c) std::swap(edge1[c], edge2[c]);
      cross(r, edge2, pvec);
      det = dot3(edge1, pvec);
      if (std::fabs(det) < TOL) continue;
      for (int c = 0; c < 3; ++c) tvec[c] = R0[c] - v0[c];
      const double u = dot3(tvec, pvec);
      cross(tvec, edge1, qvec);
      const double v = dot3(r, qvec);
      if (det > 0.0) {
        if (u < -TOL || u > det + TOL) continue;
        if (v < -TOL || u + v > det + TOL) continue;
      } else {
        if (u > TOL || u < det - TOL) continue;
        if (v > TOL || u + v < det - TOL) continue;
      }
      t = dot3(edge2, qvec) / det;
      if (Nm == 1) {
        success = true;
        break;
      }
      if (t > 1.0) continue;
      llist.push_back(t);
    }
    if (Nm == 1) break;
    if (llist.empty() || (llist.size() == 1 && std::fabs(llist[0]) < 1.0e-12)) continue;
    std::sort(llist.begin(), llist.end());
    size_t nkeep = 1;
    for (size_t i = 1; i < llist.size(); ++i)
      if (std::fabs(llist[i] - llist[nkeep - 1]) > 1.0e-12) llist[nkeep++] = llist[i];
    llist.resize(nkeep);
    for (size_t il = 0; il + 1 < nkeep; il += 2) pathlen[mat - 1] += llist[il + 1] - llist[il];
    if (nkeep % 2 == 1) pathlen[mat - 1] += 1.0 - llist[nkeep - 1];
  }
  if (Nm == 1) {
    if (!success) return false;
    pathlen[0] = (1.0 - t) * length;
  } else {
    for (int m = 0; m < Nm; ++m) pathlen[m] *= length;
  }
  return true;
}

// phi0ek [NENK]: the uncollided fluence n exp(-sum_m attn_m l_m) at the
// quadrature points in the beam (rays from the back-projection of each point on
// the source plane, RayTracing.f90:149-199), integrated against the shape
// functions and multiplied by IMr (FCSSource.f90:134-176, identical in
// AddAngularFluence.f90:109-151).  quadmask [NENK] (may be NULL) receives the
// in-beam flag of each quadrature point.  Returns -1 if a ray found no entry.
int orc_beam_uncollided(void* h, const double* axis, const double* origin, const double* e1, const double* e2,
                        int cutout, const double* params, double n, const double* attn, double* phi0ek,
                        int32_t* quadmask) {
  OrcOp* op = static_cast<OrcOp*>(h);
  OrcBeam B;
  for (int c = 0; c < 3; ++c) {
    B.axis[c] = axis[c];
    B.origin[c] = origin[c];
    B.e1[c] = e1[c];
    B.e2[c] = e2[c];
  }
  B.cutout = cutout;
  for (int c = 0; c < 4; ++c) B.params[c] = params ? params[c] : 0.0;
  B.n = n;
  std::vector<OrcRegion> reg;
  build_regions(op, reg);
  double cq[4][3], wq[4], u[4][4];
  tet_quadrature(cq, wq);
  for (int kq = 0; kq < 4; ++kq)
    for (int k = 0; k < 4; ++k) u[k][kq] = tet_sf(k, cq[kq]);
  const i64 NE = op->NE;
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : failed)
  for (i64 e = 0; e < NE; ++e) {
    double re[4][3], v[4][3], phi0q[4];
    bool mask[4];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) re[k][c] = op->rg[op->node[e * 4 + k] * 3 + c];
    for (int kq = 0; kq < 4; ++kq)   // MapToTetrahedralQuad
      for (int c = 0; c < 3; ++c) {
        double x = 0.0;
        for (int k = 0; k < 4; ++k) x += tet_sf(k, cq[kq]) * re[k][c];
        v[kq][c] = x;
      }
    std::vector<double> pl(op->Nm);
    for (int kq = 0; kq < 4; ++kq) {
      mask[kq] = in_field(B, v[kq]);
      phi0q[kq] = 0.0;
      if (!mask[kq]) continue;
      double r[3], Rp[3];
      for (int c = 0; c < 3; ++c) r[c] = v[kq][c] - B.origin[c];
      const double a1 = dot3(r, B.e1), a2 = dot3(r, B.e2);
      for (int c = 0; c < 3; ++c) Rp[c] = B.origin[c] + a1 * B.e1[c] + a2 * B.e2[c];
      for (int c = 0; c < 3; ++c) r[c] = v[kq][c] - Rp[c];
      if (!trace_vector(op, reg, Rp, r, pl.data())) {
        failed = 1;
        continue;
      }
      double tau = 0.0;   // DOT_PRODUCT(attn, pathlen)
      for (int m = 0; m < op->Nm; ++m) tau += attn[m] * pl[m];
      phi0q[kq] = B.n * (1.0 * std::exp(-tau));
    }
    double acc[4];
    for (int k = 0; k < 4; ++k) {
      acc[k] = 0.0;
      for (int kq = 0; kq < 4; ++kq)
        if (mask[kq]) acc[k] = acc[k] + wq[kq] * u[k][kq] * phi0q[kq];
    }
    for (int k = 0; k < 4; ++k) {
      double y = 0.0;
      for (int j = 0; j < 4; ++j) y += op->IMr[k * 4 + j] * acc[j];
      phi0ek[4 * e + k] = y;
    }
    if (quadmask)
      for (int kq = 0; kq < 4; ++kq) quadmask[4 * e + kq] = mask[kq] ? 1 : 0;
  }
  return failed ? -1 : 0;
}

// sum_l (2l+1) P_l(x) c_l / (4 pi), P_l by the three-term recurrence (the
// reference evaluates core's Legendre coefficients by Horner's rule).
static double legendre_series(int L, double x, const double* c) {
  double p0 = 1.0, p1 = x, f = 0.0;
  for (int l = 0; l <= L; ++l) {
    double P;
    if (l == 0) P = 1.0;
    else if (l == 1) P = x;
    else {
      P = ((2 * l - 1) * x * p1 - (l - 1) * p0) / l;
      p0 = p1;
      p1 = P;
    }
    f += (2 * l + 1) * P * c[l] / (4.0 * PI);
  }
  return f;
}

// FCSSource.f90:178-215 (planar): s(sdof + i NENK) += phi0ek(sdof) angfac(i, m),
// angfac = sum_l (2l+1) P_l(k_i . k0) sigma_l(m) / (4 pi) with the operator's physics.
void orc_beam_source(void* h, const double* axis, const double* phi0ek, double* s) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 NENK = op->NENK;
  const int L1 = op->L + 1;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < op->NI; ++i) {
    const double kk0 = dot3(&op->kdir[i * 3], axis);
    for (int m = 0; m < op->Nm; ++m) {
      const double angfac = legendre_series(op->L, kk0, &op->sigs[m * L1]);
      for (i64 sd : op->mat2sd[m]) s[sd + i * NENK] = s[sd + i * NENK] + phi0ek[sd] * angfac;
    }
  }
}

// AddAngularFluence.f90:153-166 with GenAngularDistribution (Sources.f90:257-276):
// unc(sdof + i NENK) += phi0ek(sdof) sum_l (2l+1) P_l(k_i . k0) / (4 pi).
void orc_beam_angular_fluence(void* h, const double* axis, const double* phi0ek, double* unc) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 NENK = op->NENK;
  std::vector<double> one(op->L + 1, 1.0);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < op->NI; ++i) {
    const double angfac = legendre_series(op->L, dot3(&op->kdir[i * 3], axis), one.data());
    for (i64 sd = 0; sd < NENK; ++sd) unc[sd + i * NENK] = unc[sd + i * NENK] + phi0ek[sd] * angfac;
  }
}

// SpaceAngleVector%Fluence (LinearOperatorClass.f90:154-174).
void orc_fluence(void* h, const double* v, double* fl) {
  OrcOp* op = static_cast<OrcOp*>(h);
  const i64 NENK = op->NENK;
  for (i64 j = 0; j < NENK; ++j) fl[j] = 0.0;
  for (i64 i = 0; i < op->NI; ++i)
    for (i64 j = 0; j < NENK; ++j) fl[j] = fl[j] + op->w[i] * v[j + i * NENK];
}

double orc_norm2(const double* x, i64 n) { return norm2_scaled(x, n); }

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
