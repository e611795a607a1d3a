// Host-side precompute for an operator: geometry, upwind schedule, angular matrices.
//
// Compiled with -ffp-contract=off: the upstream/outflow decisions of the schedule
// must come from exactly the arithmetic the reference writes (k.n as three
// products summed left to right, SweepOrder.f90:124-126, TransportAlgebra.f90:149),
// so that the schedule is bit-identical to the reference's ordering rule.
#include <omp.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "lb_internal.hpp"

namespace lb {

namespace {

const double kPi = 3.14159265358979323846264338327950288;
const double kEps = 2.220446049250313e-16;  // EPSILON(1d0), SweepOrder.f90:22

// Fortran NORM2 as gfortran evaluates it: scaled sum of squares.
double norm2_3(const double* x) {
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < 3; ++i) {
    if (x[i] == 0.0) continue;
    double a = std::fabs(x[i]);
    if (a > scale) {
      double r = scale / a;
      ssq = 1.0 + ssq * r * r;
      scale = a;
    } else {
      double r = a / scale;
      ssq += r * r;
    }
  }
  return scale * std::sqrt(ssq);
}

inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

}  // namespace

int upload(lb_op* op, DevBuf& dst, const void* src, size_t bytes) {
  if (op->host_only) return LB_OK;
  LB_TRY(dst.reserve(bytes));
  LB_HIP(hipMemcpy(dst.p, src, bytes, hipMemcpyHostToDevice));
  return LB_OK;
}

int compute_geometry(lb_op* op, const double* rg, const int64_t* node, const int64_t* nbr_in, const int64_t* mat,
                     int base) {
  const int64_t NE = op->NE, NK = op->NK;
  if (NE >= (int64_t(1) << 24)) {
    set_error("lb_op_create: the sweep packs element ids in 24 bits; NE must be below 16777216");
    return LB_E_ARG;
  }
  std::vector<int64_t> nd(NE * 4), nb(NE * 4);
  op->h_mat.assign(NE, 0);
  int Nm = 1;
  for (int64_t j = 0; j < NE * 4; ++j) {
    nd[j] = node[j] - base;
    nb[j] = nbr_in[j] - base;   // boundary becomes -1
    if (nd[j] < 0 || nd[j] >= NK || nb[j] < -1 || nb[j] >= NE) {
      set_error("lb_op_create: node or neighbour index out of range");
      return LB_E_ARG;
    }
  }
  for (int64_t e = 0; e < NE; ++e) {
    int64_t m = mat[e] - base;
    if (m < 0 || m >= INT_MAX) {
      set_error("lb_op_create: material index out of range");
      return LB_E_ARG;
    }
    op->h_mat[e] = (int32_t)m;
    Nm = std::max<int>(Nm, (int)m + 1);
  }
  op->Nm = Nm;
  // element vertices in local node order, for ray tracing an external beam
  op->h_vert.assign(NE * 12, 0.0);
  for (int64_t e = 0; e < NE; ++e)
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) op->h_vert[e * 12 + k * 3 + c] = rg[nd[e * 4 + k] * 3 + c];
  // the neighbour relation must be symmetric
  for (int64_t e = 0; e < NE; ++e)
    for (int f = 0; f < 4; ++f) {
      int64_t ep = nb[e * 4 + f];
      if (ep < 0) continue;
      int hits = 0;
      for (int g = 0; g < 4; ++g) hits += nb[ep * 4 + g] == e;
      if (hits != 1) {
        set_error("lb_op_create: neighbour table is not symmetric (element " + std::to_string(e) + ")");
        return LB_E_ARG;
      }
    }

  // Normals (PostProcess.f90:48-102): from the sorted global node triple of the
  // face, n = (x3 - x1) x (x2 - x1), signed away from the centroid and divided by
  // NORM2(n); the face's other element holds the exact negation.  The element with
  // the lower index computes it (ascending loop with the `done` flag).
  op->h_nrm.assign(NE * 12, 0.0);
  std::vector<double> area(NE * 4), detJ(NE);
  for (int64_t e = 0; e < NE; ++e) {
    const int64_t* en = &nd[e * 4];
    double com[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) com[c] += rg[en[k] * 3 + c];
    for (int c = 0; c < 3; ++c) com[c] /= 4.0;
    for (int f = 0; f < 4; ++f) {
      int64_t t[3] = {en[kFaceNodes[f][0]], en[kFaceNodes[f][1]], en[kFaceNodes[f][2]]};
      std::sort(t, t + 3);
      double r1[3], r2[3], n[3];
      for (int c = 0; c < 3; ++c) {
        r1[c] = rg[t[1] * 3 + c] - rg[t[0] * 3 + c];
        r2[c] = rg[t[2] * 3 + c] - rg[t[0] * 3 + c];
      }
      n[0] = r2[1] * r1[2] - r2[2] * r1[1];
      n[1] = r2[2] * r1[0] - r2[0] * r1[2];
      n[2] = r2[0] * r1[1] - r2[1] * r1[0];
      double nn = norm2_3(n);
      area[e * 4 + f] = 0.5 * nn;
      int64_t ep = nb[e * 4 + f];
      double* out = &op->h_nrm[e * 12 + f * 3];
      if (ep >= 0 && ep < e) {
        int fp = 0;
        while (nb[ep * 4 + fp] != e) ++fp;
        for (int c = 0; c < 3; ++c) out[c] = -op->h_nrm[ep * 12 + fp * 3 + c];
        continue;
      }
      double d[3] = {rg[t[0] * 3] - com[0], rg[t[0] * 3 + 1] - com[1], rg[t[0] * 3 + 2] - com[2]};
      double fac = (dot3(n, d) >= 0.0 ? 1.0 : -1.0) / nn;
      for (int c = 0; c < 3; ++c) out[c] = fac * n[c];
    }
    // detJ as Jacobians.f90:38-80 forms it
    double J[3][3];
    for (int j = 0; j < 3; ++j)
      for (int c = 0; c < 3; ++c) J[c][j] = rg[en[j + 1] * 3 + c] - rg[en[0] * 3 + c];
    detJ[e] = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (!(detJ[e] > 0.0)) {
      set_error("lb_op_create: element " + std::to_string(e) +
                " has a non-positive Jacobian; tetrahedra must be positively oriented");
      return LB_E_ARG;
    }
  }

  // Compact per-face geometry: unit normal and c_f = 10 A_f / detJ = 5 A_f / (3 V).
  // With it the mass-eliminated matrices of a linear tet are (closed forms of
  // FEAInnerProducts.f90:311-407):  -sum_d k_d IC_d = c_{F(a)} (k.n_{F(a)}) in every
  // column of row a (F(a) = face opposite node a), IFM_f(a,b) = c_f (delta_ab +
  // [a in f] - 4/5) for b in f, and IIM_f the same with columns on the neighbour's nodes.
  std::vector<double> geo(NE * 16);
  std::vector<uint32_t> perm(NE, 0);
  std::vector<int32_t> nbr32(NE * 4);
  for (int64_t e = 0; e < NE; ++e) {
    for (int f = 0; f < 4; ++f) {
      for (int c = 0; c < 3; ++c) geo[e * 16 + f * 4 + c] = op->h_nrm[e * 12 + f * 3 + c];
      geo[e * 16 + f * 4 + 3] = 10.0 * area[e * 4 + f] / detJ[e];
      int64_t ep = nb[e * 4 + f];
      nbr32[e * 4 + f] = (int32_t)ep;
      if (ep < 0) continue;
      for (int j = 0; j < 3; ++j) {
        int64_t gnode = nd[e * 4 + kFaceNodes[f][j]];
        int kp = 0;
        while (kp < 4 && nd[ep * 4 + kp] != gnode) ++kp;
        if (kp == 4) {
          set_error("lb_op_create: neighbouring elements " + std::to_string(e) + " and " + std::to_string(ep) +
                    " do not share the face nodes");
          return LB_E_ARG;
        }
        perm[e] |= uint32_t(kp) << (6 * f + 2 * j);
      }
    }
  }
  LB_TRY(upload(op, op->d_geo, geo.data(), geo.size() * sizeof(double)));
  LB_TRY(upload(op, op->d_nbr, nbr32.data(), nbr32.size() * sizeof(int32_t)));
  op->h_nbr = std::move(nbr32);
  op->h_detJ = detJ;
  LB_TRY(upload(op, op->d_perm, perm.data(), perm.size() * sizeof(uint32_t)));
  LB_TRY(upload(op, op->d_mat, op->h_mat.data(), op->h_mat.size() * sizeof(int32_t)));
  return compute_schedule(op, nb.data());
}

// Upwind schedule.  SweepOrder.f90:105-207 peels levels by repeated passes: a
// pass takes, in ascending element order, every element with no remaining
// upstream dependency, and clears dependencies only after the pass.  The result
// is "levels concatenated, ascending element index within a level" where the
// level of an element is one more than the deepest of its upstream neighbours.
// Here the same order is produced in O(NE) per direction with Kahn's algorithm on
// in-degree counters, processing one whole frontier (level) at a time.
int compute_schedule(lb_op* op, const int64_t* nb) {
  const int64_t NE = op->NE, NI = op->NI;
  // local face of the neighbour that faces back (findloc in SweepOrder.f90:169)
  std::vector<int8_t> back(NE * 4, -1);
  for (int64_t e = 0; e < NE; ++e)
    for (int f = 0; f < 4; ++f) {
      int64_t ep = nb[e * 4 + f];
      if (ep < 0) continue;
      int g = 0;
      while (nb[ep * 4 + g] != e) ++g;
      back[e * 4 + f] = (int8_t)g;
    }
  op->h_tasks.assign(NI * NE, 0);
  op->h_nlev.assign(NI, 0);
  std::vector<std::vector<int32_t>> lvl(NI);
  int stagnant = 0;
#pragma omp parallel
  {
    std::vector<uint8_t> upm(NE), outm(NE), indeg(NE);
    std::vector<int32_t> cur, nxt;
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < NI; ++i) {
      const double* k = &op->h_k[i * 3];
      cur.clear();
      for (int64_t e = 0; e < NE; ++e) {
        uint8_t u = 0, o = 0;
        for (int f = 0; f < 4; ++f) {
          double kn = dot3(k, &op->h_nrm[e * 12 + f * 3]);
          if (nb[e * 4 + f] >= 0 && kn < -kEps) u |= uint8_t(1u << f);
          if (!(kn < 0.0)) o |= uint8_t(1u << f);   // TransportAlgebra.f90:150 keeps kn >= 0
        }
        upm[e] = u;
        outm[e] = o;
        indeg[e] = (uint8_t)__builtin_popcount(u);
        if (indeg[e] == 0) cur.push_back((int32_t)e);
      }
      uint32_t* tk = &op->h_tasks[i * NE];
      std::vector<int32_t>& off = lvl[i];
      off.clear();
      int64_t es = 0;
      while (!cur.empty()) {
        off.push_back((int32_t)es);
        nxt.clear();
        for (int32_t e : cur) {
          tk[es++] = uint32_t(e) | (uint32_t(upm[e]) << 24) | (uint32_t(outm[e]) << 28);
          for (int f = 0; f < 4; ++f) {
            int64_t ep = nb[e * 4 + f];
            if (ep < 0) continue;
            int g = back[e * 4 + f];
            if (upm[ep] >> g & 1) {
              if (--indeg[ep] == 0) nxt.push_back((int32_t)ep);
            }
          }
        }
        std::sort(nxt.begin(), nxt.end());
        cur.swap(nxt);
      }
      off.push_back((int32_t)es);
      if (es != NE) {
#pragma omp atomic write
        stagnant = 1;
      }
      op->h_nlev[i] = (int32_t)off.size() - 1;
    }
  }
  if (stagnant) {
    set_error("SweepOrder: Sweep is stagnant. The upwind dependency graph of some direction has a cycle "
              "or the mesh connectivity is inconsistent.");
    return LB_E_STAGNANT;
  }
  op->maxlev = 0;
  for (int64_t i = 0; i < NI; ++i) op->maxlev = std::max(op->maxlev, op->h_nlev[i]);
  op->h_lvl_off.assign(NI * (op->maxlev + 1), 0);
  for (int64_t i = 0; i < NI; ++i) {
    int32_t* o = &op->h_lvl_off[i * (op->maxlev + 1)];
    for (int l = 0; l <= op->maxlev; ++l) o[l] = l < (int)lvl[i].size() ? lvl[i][l] : (int32_t)NE;
  }
  if (op->host_only) return LB_OK;
  // The device sweeps each level in its own order: the cells whose upstream
  // neighbours all sit in the last `slots` positions before the level's end
  // come first, the others after them, each part in the reference's ascending
  // element order.  Recent producers are the ones the LDS sweep's slot plan
  // keeps in LDS, so this groups most of the cells that read a trace from
  // global memory at the end of their level, and most waves take one path.
  // Cells of a level are independent, so every order gives the same flux;
  // h_tasks keeps the reference order for the schedule queries.
  int dev = 0, lds = 0;
  LB_HIP(hipGetDevice(&dev));
  LB_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  op->lds_max = lds;
  int64_t slots = 1;
  while (slots * 2 * 4 * (int64_t)sizeof(double) <= lds) slots *= 2;
  std::vector<uint32_t> dev_tasks(NI * NE);
#pragma omp parallel
  {
    std::vector<int32_t> fpos(NE);
    std::vector<uint32_t> miss;
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < NI; ++i) {
      const uint32_t* tk = &op->h_tasks[i * NE];
      uint32_t* dk = &dev_tasks[i * NE];
      const int32_t* off = &op->h_lvl_off[i * (op->maxlev + 1)];
      for (int l = 0; l < op->h_nlev[i]; ++l) {
        const int64_t s0 = off[l], s1 = off[l + 1], lim = s1 - slots;
        int64_t o = s0;
        miss.clear();
        for (int64_t sp = s0; sp < s1; ++sp) {
          const uint32_t w = tk[sp];
          const int64_t e = w & 0xFFFFFFu;
          const unsigned up = (w >> 24) & 0xFu;
          bool hit = true;
          for (int f = 0; f < 4; ++f)
            if ((up >> f & 1u) && fpos[nb[e * 4 + f]] < lim) hit = false;
          if (hit) {
            dk[o] = w;
            fpos[e] = (int32_t)o++;
          } else {
            miss.push_back(w);
          }
        }
        for (uint32_t w : miss) {
          dk[o] = w;
          fpos[w & 0xFFFFFFu] = (int32_t)o++;
        }
      }
    }
  }
  LB_TRY(upload(op, op->d_tasks, dev_tasks.data(), dev_tasks.size() * sizeof(uint32_t)));
  return LB_OK;
}

// Angular matrices: YY (SNPrep.f90:106-133) for the dense kernel and the real
// spherical-harmonic moment matrices D (w_i Y_lm(k_i)) and M (Y_lm(k_i)), which
// factor it exactly: sum_m Y_lm(k) Y_lm(k') = (2l+1)/(4 pi) P_l(k.k').
void real_sph_harm(int L, const double* k, double* y) {
  double mu = std::max(-1.0, std::min(1.0, k[2]));
  double s = std::sqrt(std::max(0.0, 1.0 - mu * mu));
  double phi = std::atan2(k[1], k[0]);
  // associated Legendre P_l^m(mu) without the Condon-Shortley phase, P[l*(L+1)+m]
  std::vector<double> P((L + 1) * (L + 1), 0.0);
  P[0] = 1.0;
  for (int m = 1; m <= L; ++m) P[m * (L + 1) + m] = (2 * m - 1) * s * P[(m - 1) * (L + 1) + m - 1];
  for (int m = 0; m < L; ++m) P[(m + 1) * (L + 1) + m] = (2 * m + 1) * mu * P[m * (L + 1) + m];
  for (int m = 0; m <= L; ++m)
    for (int l = m + 2; l <= L; ++l)
      P[l * (L + 1) + m] = ((2 * l - 1) * mu * P[(l - 1) * (L + 1) + m] - (l + m - 1) * P[(l - 2) * (L + 1) + m]) / (l - m);
  for (int l = 0; l <= L; ++l)
    for (int m = -l; m <= l; ++m) {
      int am = std::abs(m);
      double ratio = 1.0;  // (l-|m|)! / (l+|m|)!
      for (int t = l - am + 1; t <= l + am; ++t) ratio /= t;
      double nrm = std::sqrt((2 * l + 1) / (4.0 * kPi) * ratio);
      double v;
      if (m == 0) v = nrm * P[l * (L + 1)];
      else if (m > 0) v = std::sqrt(2.0) * nrm * P[l * (L + 1) + am] * std::cos(am * phi);
      else v = std::sqrt(2.0) * nrm * P[l * (L + 1) + am] * std::sin(am * phi);
      y[l * l + l + m] = v;
    }
}

int build_angular(lb_op* op) {
  const int64_t NI = op->NI;
  const int L = op->L;
  op->nmom = (L + 1) * (L + 1);
  op->h_yy.assign((size_t)(L + 1) * NI * NI, 0.0);
  for (int64_t ip = 0; ip < NI; ++ip)
    for (int64_t i = 0; i < NI; ++i) {
      double x = (i == ip) ? 1.0 : dot3(&op->h_k[i * 3], &op->h_k[ip * 3]);
      double pm2 = 1.0, pm1 = x;
      for (int l = 0; l <= L; ++l) {
        double P = (l == 0) ? 1.0 : (l == 1 ? x : ((2 * l - 1) * x * pm1 - (l - 1) * pm2) / l);
        if (l >= 2) {
          pm2 = pm1;
          pm1 = P;
        }
        op->h_yy[((size_t)l * NI + ip) * NI + i] = op->h_w[ip] * (2 * l + 1) * P / (4.0 * kPi);
      }
    }
  const int nm = op->nmom;
  op->h_D.assign((size_t)nm * NI, 0.0);
  op->h_M.assign((size_t)NI * nm, 0.0);
  op->h_deg.assign(nm, 0);
  std::vector<double> y(nm);
  for (int l = 0; l <= L; ++l)
    for (int m = -l; m <= l; ++m) op->h_deg[l * l + l + m] = l;
  for (int64_t i = 0; i < NI; ++i) {
    real_sph_harm(L, &op->h_k[i * 3], y.data());
    for (int r = 0; r < nm; ++r) {
      op->h_M[i * nm + r] = y[r];
      op->h_D[(size_t)r * NI + i] = op->h_w[i] * y[r];
    }
  }
  LB_TRY(upload(op, op->d_D, op->h_D.data(), op->h_D.size() * sizeof(double)));
  LB_TRY(upload(op, op->d_M, op->h_M.data(), op->h_M.size() * sizeof(double)));
  LB_TRY(upload(op, op->d_deg, op->h_deg.data(), op->h_deg.size() * sizeof(int32_t)));
  LB_TRY(upload(op, op->d_k, op->h_k.data(), op->h_k.size() * sizeof(double)));
  LB_TRY(upload(op, op->d_w, op->h_w.data(), op->h_w.size() * sizeof(double)));
  return LB_OK;
}

// Work list for sweeping G groups at once: each workgroup of 256 lanes takes
// cpw = 256 / G consecutive cells of one level of one direction, one lane per
// (cell, group).  Entries of a global level are ordered so that consecutive
// workgroup ids walk different directions, spreading each level's directions
// over the XCDs.
LevelTable* level_table(lb_op* op, int G, int* rc) {
  int cpw = std::max(1, 256 / G);
  auto it = op->tables.find(cpw);
  if (it != op->tables.end()) {
    *rc = LB_OK;
    return &it->second;
  }
  LevelTable t;
  t.cpw = cpw;
  const int64_t NI = op->NI;
  t.level_off.assign(op->maxlev + 1, 0);
  for (int l = 0; l < op->maxlev; ++l) {
    t.level_off[l] = (int32_t)(t.entries.size() / 4);
    for (int64_t i = 0; i < NI; ++i) {
      const int32_t* o = &op->h_lvl_off[i * (op->maxlev + 1)];
      if (l >= op->h_nlev[i]) continue;
      for (int32_t s = o[l]; s < o[l + 1]; s += cpw) {
        t.entries.push_back((int32_t)i);
        t.entries.push_back(s);
        t.entries.push_back(std::min(cpw, o[l + 1] - s));
        t.entries.push_back(0);
      }
    }
  }
  t.level_off[op->maxlev] = (int32_t)(t.entries.size() / 4);
  *rc = t.d_entries.reserve(t.entries.size() * sizeof(int32_t));
  if (*rc != LB_OK) return nullptr;
  hipError_t e = hipMemcpy(t.d_entries.p, t.entries.data(), t.entries.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("HIP error ") + hipGetErrorString(e) + " uploading the sweep work list");
    *rc = LB_E_HIP;
    return nullptr;
  }
  auto res = op->tables.emplace(cpw, std::move(t));
  *rc = LB_OK;
  return &res.first->second;
}

}  // namespace lb
