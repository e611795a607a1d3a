// gfx950 kernels of the transport-solve path.
//
// Sweep: level-synchronous upwind sweep of TransportAlgebra.f90:29-246.  One
// launch per global level sweeps that level of every direction at once; a
// 256-lane workgroup takes consecutive cells of one level of one direction, one
// lane per (cell, group).  Each lane gathers the three face-node values of its
// upstream neighbours, builds the 4x4 local matrix in registers from the compact
// face data (unit normal + c_f), and solves it with the reference's partial-pivot
// LU.  Vector layout: v[((i*NE + e)*4 + k)*G + g] (groups fastest) so the lanes of
// one cell read consecutive doubles.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstring>

#include "lb_internal.hpp"

namespace lb {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ bool in_face(int f, int a) {
  // node a lies on local face f unless it is the opposite node 3 - f
  return a != 3 - f;
}

// Per-(direction, sweep position) record, 48 bytes, in sweep order so that the
// lanes of one level read consecutive records.  ck holds c_f (k . n_f) of faces
// 0-2: every term of the local matrix and of the upstream coupling of a linear
// tet is a multiple of one of the four (closed forms of
// FEAInnerProducts.f90:311-407), and the fourth is minus the sum of the three,
// since sum_f c_f n_f = (10 / detJ) sum_f A_f n_f = 0 over a closed tetrahedron.
struct TaskRec {
  uint32_t w;      // cell | upstream mask << 24 | outflow mask << 28
  uint32_t pm;     // neighbour local node of each face node, 2 bits each (24 bits) | material << 24
  int32_t nb[4];   // neighbour across each local face, -1 on the boundary
  double ck[3];    // c_f (k . n_f) of faces 0, 1, 2
};
static_assert(sizeof(TaskRec) == 48, "three 16-byte loads per record");

// Kinds of LDS trace handling of the walks: none; LDS slots by the slot plan;
// the same with the one-group FP32 flux output (kLdsF32): the flux goes to an
// FP32 vector in sweep order, and the FP64 flux is written only for the cells
// whose flux some consumer reads from global memory (the plan's gprod marks).
constexpr int kLdsNone = 0, kLdsPlan = 1, kLdsF32 = 2;

// Arguments shared by the sweep kernels.
struct SweepArgs {
  const TaskRec* __restrict__ rec;
  const double* __restrict__ sigt;
  int Gtot, g0, G;
  int64_t NE;
  const double* q;   // may alias psi (in-place sweep)
  double* psi;
  int vec1 = 0;      // one group and a 16-byte-aligned psi: 16-byte flux stores
  int pos_out = 0;   // flux stored at the cell's sweep position instead of its element (one group); the
                     // records then carry the neighbours' positions and the staging entries positions
  int pos_in = 0;    // source read at the cell's sweep position (with pos_out; q may then alias psi)
  float* psi32 = nullptr;                        // kLdsF32: the flux in FP32, sweep order
  const uint8_t* __restrict__ gprod = nullptr;   // kLdsF32: [dir][pos] 1 = the FP64 flux is read globally
  // LDS sweep only: the slot plan of this group-batch size (lb_trace_plan.cpp)
  const short4* __restrict__ uslot = nullptr;    // [dir][pos]: LDS slot of each upstream trace, -1 global
  const int16_t* __restrict__ wslot = nullptr;   // [dir][pos]: slot the cell writes, -1 none
  const int2* __restrict__ stage = nullptr;      // (LDS slot, element) copied from global memory
};

// c_f (k . n_f) of one face, with contraction off: the same four products and
// two sums whatever the compiler would fuse in its surroundings.
__device__ __forceinline__ double face_ck(const double4& gf, double kx, double ky, double kz) {
#pragma clang fp contract(off)
  return gf.w * (kx * gf.x + ky * gf.y + kz * gf.z);
}

// LDS of the direction-resident LDS sweep: trace slots laid out [slot][node][group].
// Which cell writes which slot, and where each upstream trace is read, is fixed
// by the slot plan (lb_trace_plan.cpp): ring slots held from the producer's level
// until its last consumer, and staging slots filled from global memory one level
// ahead.
struct TraceRing {
  double* buf;
};

// With many groups per cell the records and sources are read once per sweep and
// share no cache lines between cells; marking them non-temporal keeps L2 for the
// upstream face traces (measured: 24 groups 53.8 -> 50.6 ms per sweep; with one
// group, where neighbouring cells share source lines, it costs instead).
constexpr int kStreamGroups = 8;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double dbl_of(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The part of a cell's solve that does not depend on upstream results.  Kernels
// may fetch it one level ahead.
struct CellPrep {
  uint32_t w;
  double ck[4];
  double sig;
  double b[4];
  int4 nb;
  uint32_t pm;  // node permutation (24 bits)
  int ws;     // LDS sweep only: slot this cell writes and slots of its upstream traces
  int4 us;
  int gp;     // kLdsF32: the FP64 flux is also written to global memory
};

// A cell's record as loaded (and, in the LDS sweep, its slot-plan entries);
// decode_rec turns it into the CellPrep and issues the loads that depend on it.
struct RecRaw {
  u32x4 a, b, c;
  int ws;
  short4 us;
  int gp;
};

template <bool NT, int LDS>
__device__ __forceinline__ void load_rec(const SweepArgs& A, int i, int s, RecRaw& R) {
  const int64_t k = (int64_t)i * A.NE + s;
  const u32x4* q = reinterpret_cast<const u32x4*>(A.rec + k);
  if constexpr (NT) {
    R.a = __builtin_nontemporal_load(q);
    R.b = __builtin_nontemporal_load(q + 1);
    R.c = __builtin_nontemporal_load(q + 2);
  } else {
    R.a = q[0];
    R.b = q[1];
    R.c = q[2];
  }
  if constexpr (LDS) {
    R.ws = A.wslot[k];
    R.us = A.uslot[k];
  }
  if constexpr (LDS == kLdsF32) R.gp = A.gprod[k];
}

template <bool NT, int LDS>
__device__ __forceinline__ void decode_rec(const SweepArgs& A, int i, int s, int g, const RecRaw& R, CellPrep& P) {
  if constexpr (LDS) {
    P.ws = R.ws;
    P.us = make_int4(R.us.x, R.us.y, R.us.z, R.us.w);
  }
  if constexpr (LDS == kLdsF32) P.gp = R.gp;
  P.w = R.a.x;
  P.pm = R.a.y & 0xFFFFFFu;
  P.nb = make_int4((int)R.a.z, (int)R.a.w, (int)R.b.x, (int)R.b.y);
  P.ck[0] = dbl_of(R.b.z, R.b.w);
  P.ck[1] = dbl_of(R.c.x, R.c.y);
  P.ck[2] = dbl_of(R.c.z, R.c.w);
  P.ck[3] = -((P.ck[0] + P.ck[1]) + P.ck[2]);
  P.sig = A.sigt[(int64_t)(R.a.y >> 24) * A.Gtot + A.g0 + g];
  const int src = A.pos_in ? s : int(P.w & 0xFFFFFFu);   // where the source is
  const int64_t base = ((int64_t)i * A.NE + src) * 4 * A.G + g;
#pragma unroll
  for (int a = 0; a < 4; ++a) P.b[a] = NT ? __builtin_nontemporal_load(A.q + base + a * A.G) : A.q[base + a * A.G];
}

template <bool NT, int LDS = 0>
__device__ __forceinline__ void prep_cell(const SweepArgs& A, int i, int s, int g, CellPrep& P) {
  RecRaw R;
  load_rec<NT, LDS>(A, i, s, R);
  decode_rec<NT, LDS>(A, i, s, g, R, P);
}

// Where a cell reads its upstream traces and what it writes besides the flux.
enum TraceSrc { kTraceGlobal = 0, kTraceGlobalRing = 1, kTraceRing = 2 };

// Upstream face traces of a cell: u[f][j] = the neighbour's flux at the j-th node
// of local face f, for the upstream faces only (from the ring or global memory).
template <int SRC>
__device__ __forceinline__ void load_upstream(const SweepArgs& A, int i, const CellPrep& P, int g,
                                              const TraceRing& ring, double (&u)[4][3]) {
  const unsigned up = (P.w >> 24) & 0xFu;
  if (up) {
    const int nbv[4] = {P.nb.x, P.nb.y, P.nb.z, P.nb.w};
    int usv[4] = {0, 0, 0, 0};
    if constexpr (SRC == kTraceRing) {
      usv[0] = P.us.x;
      usv[1] = P.us.y;
      usv[2] = P.us.z;
      usv[3] = P.us.w;
    }
    const uint32_t pm = P.pm;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (up >> f & 1u) {
        const unsigned n0 = (pm >> (6 * f)) & 3u, n1 = (pm >> (6 * f + 2)) & 3u, n2 = (pm >> (6 * f + 4)) & 3u;
        if constexpr (SRC == kTraceRing) {
          const double* rb = ring.buf + (int64_t)usv[f] * 4 * A.G + g;
          u[f][0] = rb[n0 * A.G];
          u[f][1] = rb[n1 * A.G];
          u[f][2] = rb[n2 * A.G];
        } else {
          const int64_t ub = ((int64_t)i * A.NE + nbv[f]) * 4 * A.G + g;
          u[f][0] = A.psi[ub + n0 * A.G];
          u[f][1] = A.psi[ub + n1 * A.G];
          u[f][2] = A.psi[ub + n2 * A.G];
        }
      }
    }
  }
}

// The rest of a cell's solve once its traces are loaded: upstream coupling,
// local matrix, LU solve, store.
// 1 / x for a pivot: the hardware reciprocal refined by two Newton steps
// (r <- r + r (1 - x r)), a shorter dependent chain than the IEEE division.
__device__ __forceinline__ double pivot_recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

template <int SRC, int LDS = kLdsNone>
__device__ __forceinline__ void solve_with_upstream(const SweepArgs& A, int i, int s, const CellPrep& P, int g,
                                                    const TraceRing& ring, const double (&u)[4][3]) {
  const uint32_t w = P.w;
  const int dst = A.pos_out ? s : int(w & 0xFFFFFFu);   // where the flux goes
  const unsigned up = (w >> 24) & 0xFu;
  const unsigned out = w >> 28;
  double b[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) b[a] = P.b[a];
  const int64_t base = ((int64_t)i * A.NE + dst) * 4 * A.G + g;

  // Upstream faces (TransportAlgebra.f90:82-94): s_e += -(k.n) IIM_f psi_up with
  // IIM_f psi_up = c_f (u_a + U/5) on the face nodes and -4/5 c_f U on the opposite
  // node, u = neighbour values at the shared nodes, U their sum.
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (up >> f & 1u) {
      const double u0 = u[f][0], u1 = u[f][1], u2 = u[f][2];
      const double U = u0 + u1 + u2;
      const double coef = -P.ck[f];
      b[kFaceNodes[f][0]] += coef * (u0 + 0.2 * U);
      b[kFaceNodes[f][1]] += coef * (u1 + 0.2 * U);
      b[kFaceNodes[f][2]] += coef * (u2 + 0.2 * U);
      b[3 - f] += coef * (-0.8 * U);
    }
  }

  // Local matrix (ApplyInvertedDiagonal, TransportAlgebra.f90:126-152):
  // sigma I - sum_d k_d IC_d + sum_{f: k.n >= 0} (k.n) IFM_f.
  double T[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double r = P.ck[3 - a];
#pragma unroll
    for (int bb = 0; bb < 4; ++bb) T[a][bb] = r;
    T[a][a] += P.sig;
  }
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (out >> f & 1u) {
      const double v = P.ck[f];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int bb = kFaceNodes[f][j];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const double pat = (a == bb) ? 1.2 : (in_face(f, a) ? 0.2 : -0.8);
          T[a][bb] += v * pat;
        }
      }
    }
  }

  // LocalLU (TransportAlgebra.f90:173-246): partial pivoting with a row swap when a
  // strictly larger |A(i,k)| is found, elimination, back substitution.  Columns left
  // of the pivot are not needed afterwards and are not swapped; each pivot is
  // inverted once and reused by the elimination and the back substitution.
  double inv[4];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double mx = fabs(T[k][k]);
    int p = k;
#pragma unroll
    for (int r = k + 1; r < 4; ++r) {
      const double v = fabs(T[r][k]);
      if (v > mx) {
        mx = v;
        p = r;
      }
    }
#pragma unroll
    for (int r = k + 1; r < 4; ++r) {
      if (p == r) {
#pragma unroll
        for (int j = k; j < 4; ++j) {
          const double tmp = T[k][j];
          T[k][j] = T[r][j];
          T[r][j] = tmp;
        }
        const double tb = b[k];
        b[k] = b[r];
        b[r] = tb;
      }
    }
    inv[k] = pivot_recip(T[k][k]);
#pragma unroll
    for (int r = k + 1; r < 4; ++r) {
      const double factor = T[r][k] * inv[k];
#pragma unroll
      for (int j = k + 1; j < 4; ++j) T[r][j] = T[r][j] - factor * T[k][j];
      b[r] = b[r] - factor * b[k];
    }
  }
  inv[3] = pivot_recip(T[3][3]);
  double x[4];
  x[3] = b[3] * inv[3];
#pragma unroll
  for (int r = 2; r >= 0; --r) {
    double sacc = b[r];
#pragma unroll
    for (int j = r + 1; j < 4; ++j) sacc = sacc - T[r][j] * x[j];
    x[r] = sacc * inv[r];
  }
  // With one group a cell's four values are 32 contiguous bytes: two 16-byte
  // stores instead of four 8-byte ones (6.63 -> 6.09 ms per sweep of 384k tets
  // x 256 directions; the vector-memory counter retires stores and loads in
  // order, so fewer store instructions also shorten the look-ahead's waits).
  // The same for the source loads and the LDS writes measured no gain (loads:
  // 6.48 ms).
  if constexpr (LDS == kLdsF32) {
    // one group, sweep order: the FP32 flux for the moments, the FP64 flux only
    // where a later cell reads it from global memory (its LDS traces are FP64)
    *reinterpret_cast<float4*>(A.psi32 + base) = make_float4((float)x[0], (float)x[1], (float)x[2], (float)x[3]);
    if (P.gp) {
      double2* o = reinterpret_cast<double2*>(A.psi + base);
      o[0] = make_double2(x[0], x[1]);
      o[1] = make_double2(x[2], x[3]);
    }
  } else if (A.vec1) {
    double2* o = reinterpret_cast<double2*>(A.psi + base);
    o[0] = make_double2(x[0], x[1]);
    o[1] = make_double2(x[2], x[3]);
  } else {
#pragma unroll
    for (int a = 0; a < 4; ++a) A.psi[base + a * A.G] = x[a];
  }
  if constexpr (SRC != kTraceGlobal) {
    if (P.ws >= 0) {
      double* rb = ring.buf + (int64_t)P.ws * 4 * A.G + g;
#pragma unroll
      for (int a = 0; a < 4; ++a) rb[a * A.G] = x[a];
    }
  }
}

template <int SRC, int LDS = kLdsNone>
__device__ __forceinline__ void finish_cell_body(const SweepArgs& A, int i, int s, const CellPrep& P, int g,
                                                 const TraceRing& ring) {
  double u[4][3];
  load_upstream<SRC>(A, i, P, g, ring, u);
  solve_with_upstream<SRC, LDS>(A, i, s, P, g, ring, u);
}

__device__ __forceinline__ bool ring_hit(const CellPrep& P) {
  const unsigned up = (P.w >> 24) & 0xFu;
  return (!(up & 1u) || P.us.x >= 0) && (!(up & 2u) || P.us.y >= 0) && (!(up & 4u) || P.us.z >= 0) &&
         (!(up & 8u) || P.us.w >= 0);
}

// In the LDS sweep, a cell whose upstream traces all sit in LDS slots reads them
// from LDS only; any other reads them all from global memory.  The two paths are
// separate to the end of the cell, so a wave whose cells all hit never waits on
// the vector-memory counter, which would also wait for the look-ahead loads
// issued before it.
template <int LDS = 0>
__device__ __forceinline__ void finish_cell(const SweepArgs& A, int i, int s, const CellPrep& P, int g,
                                            const TraceRing& ring = TraceRing{nullptr}) {
  if constexpr (LDS) {
    if (ring_hit(P))
      finish_cell_body<kTraceRing, LDS>(A, i, s, P, g, ring);
    else
      finish_cell_body<kTraceGlobalRing, LDS>(A, i, s, P, g, ring);
  } else {
    finish_cell_body<kTraceGlobal>(A, i, s, P, g, ring);
  }
}

template <bool NT, int LDS = 0>
__device__ __forceinline__ void solve_cell(const SweepArgs& A, int i, int s, int g,
                                           const TraceRing& ring = TraceRing{nullptr}) {
  CellPrep P;
  prep_cell<NT, LDS>(A, i, s, g, P);
  finish_cell<LDS>(A, i, s, P, g, ring);
}

// The record of every (direction, sweep position); materials fit 8 bits
// (LB_MAX_MATERIALS, checked at creation).  With inv (sweep position of each
// element per direction) the neighbours are given by their positions, for
// sweeps that keep the flux in sweep order.
__global__ __launch_bounds__(kBlock) void build_records_kernel(int NI, int64_t NE, const uint32_t* __restrict__ tasks,
                                                               const double4* __restrict__ geo,
                                                               const int4* __restrict__ nbr,
                                                               const uint32_t* __restrict__ perm,
                                                               const int32_t* __restrict__ mat,
                                                               const double* __restrict__ kdir,
                                                               const int32_t* __restrict__ inv,
                                                               TaskRec* __restrict__ rec) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= (int64_t)NI * NE) return;
  const int i = int(k / NE);
  const uint32_t w = tasks[k];
  const int e = int(w & 0xFFFFFFu);
  TaskRec r;
  r.w = w;
  r.pm = (perm[e] & 0xFFFFFFu) | ((uint32_t)mat[e] << 24);
  const int4 nb = nbr[e];
  const int nbv[4] = {nb.x, nb.y, nb.z, nb.w};
  const int32_t* ip = inv ? inv + (int64_t)i * NE : nullptr;
#pragma unroll
  for (int f = 0; f < 4; ++f) r.nb[f] = (ip && nbv[f] >= 0) ? ip[nbv[f]] : nbv[f];
  const double kx = kdir[3 * i], ky = kdir[3 * i + 1], kz = kdir[3 * i + 2];
#pragma unroll
  for (int f = 0; f < 3; ++f) r.ck[f] = face_ck(geo[(int64_t)e * 4 + f], kx, ky, kz);
  rec[k] = r;
}

template <bool NT>
__global__ __launch_bounds__(kBlock) void sweep_level_kernel(const int4* __restrict__ work, SweepArgs A) {
  const int4 wd = work[blockIdx.x];
  const int c = threadIdx.x / A.G;
  const int g = threadIdx.x - c * A.G;
  if (c >= wd.z) return;
  solve_cell<NT>(A, wd.x, wd.y + c, g);
}

// Static part of the cells this lane sweeps in one level: up to R rounds of
// BLOCK / G cells each.
template <int R, bool NT, int LDS = 0>
__device__ __forceinline__ void prep_level(const SweepArgs& A, int i, int s0, int s1, int cpr, int c, int g,
                                           bool lane_ok, CellPrep (&P)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int s = s0 + c + r * cpr;
    if (lane_ok && s < s1) prep_cell<NT, LDS>(A, i, s, g, P[r]);
  }
}

template <int R, bool NT, int LDS = 0>
__device__ __forceinline__ void finish_level(const SweepArgs& A, int i, int s0, int s1, int cpr, int c, int g,
                                             bool lane_ok, const CellPrep (&P)[R],
                                             const TraceRing& ring = TraceRing{nullptr}) {
  if (!lane_ok) return;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int s = s0 + c + r * cpr;
    if (s < s1) finish_cell<LDS>(A, i, s, P[r], g, ring);
  }
  for (int s = s0 + c + R * cpr; s < s1; s += cpr) solve_cell<NT, LDS>(A, i, s, g, ring);   // beyond R rounds
}

// Staging copy of the LDS sweep: the traces that the cells of level l read from
// staging slots (plan entries so[l] .. so[l+1]) are loaded from global memory
// while level l - 1 is solved and written to LDS before the barrier that ends
// it.  Their producers lie two or more levels back, so the barrier before level
// l - 1 has ordered their global stores.  One double per lane: the plan keeps
// entries x 4 G within the smallest workgroup.
struct StageCopy {
  double v;
  int dst;   // LDS double index, -1 when the lane copies nothing
};

template <int LDS>
__device__ __forceinline__ StageCopy stage_load(const SweepArgs& A, int i, const int32_t* __restrict__ so, int l) {
  StageCopy c{0.0, -1};
  if constexpr (LDS) {
    if (l < 0) return c;   // not the last chunk of a level
    const int e0 = so[l], per = 4 * A.G;
    const int t = threadIdx.x;
    if (t < (so[l + 1] - e0) * per) {
      const int j = t / per, r = t - j * per;
      const int2 en = A.stage[e0 + j];
      c.dst = en.x * per + r;
      c.v = A.psi[((int64_t)i * A.NE + en.y) * per + r];
    }
  }
  return c;
}

template <int LDS>
__device__ __forceinline__ void stage_store(const TraceRing& ring, const StageCopy& c) {
  if constexpr (LDS) {
    if (c.dst >= 0) ring.buf[c.dst] = c.v;
  }
}

// Walks the levels of one direction (positions off[0..nl]) with the whole
// workgroup when each cell carries many groups (kStreamGroups or more): BLOCK / G
// cells per round, one step per level, a workgroup barrier between levels.  The
// static data of the next level's first R rounds (records, c_f k.n, cross
// section, source) is loaded into the second of two register buffers before the
// current level is solved; rounds beyond R are loaded and solved in place
// (measured: 24 groups at 1024 threads 56.4 ms per sweep, against 80.7 ms in
// chunks of R rounds, which spill more).  so: the direction's staging offsets
// (LDS sweep only).  Few groups per cell take walk_levels_two_stage.
template <int BLOCK, int R, bool NT, int LDS = 0>
__device__ __forceinline__ void walk_levels(const SweepArgs& A, int i, const int32_t* __restrict__ off, int nl,
                                            TraceRing ring = TraceRing{nullptr},
                                            const int32_t* __restrict__ so = nullptr) {
  static_assert(NT, "few groups per cell take walk_levels_two_stage");
  const int cpr = BLOCK / A.G;   // cells per round
  const int c = threadIdx.x / A.G;
  const int g = threadIdx.x - c * A.G;
  const bool lane_ok = c < cpr;
  CellPrep Pa[R], Pb[R];
  if (nl <= 0) return;
  prep_level<R, NT, LDS>(A, i, off[0], off[1], cpr, c, g, lane_ok, Pa);
  for (int l = 0; l < nl; l += 2) {
    if (l + 1 < nl) prep_level<R, NT, LDS>(A, i, off[l + 1], off[l + 2], cpr, c, g, lane_ok, Pb);
    const StageCopy sa = stage_load<LDS>(A, i, so, l + 1);
    finish_level<R, NT, LDS>(A, i, off[l], off[l + 1], cpr, c, g, lane_ok, Pa, ring);
    stage_store<LDS>(ring, sa);
    __syncthreads();
    if (l + 1 >= nl) break;
    if (l + 2 < nl) prep_level<R, NT, LDS>(A, i, off[l + 2], off[l + 3], cpr, c, g, lane_ok, Pa);
    const StageCopy sb = stage_load<LDS>(A, i, so, l + 2);
    finish_level<R, NT, LDS>(A, i, off[l + 1], off[l + 2], cpr, c, g, lane_ok, Pb, ring);
    stage_store<LDS>(ring, sb);
    __syncthreads();
  }
}

// Cells [s0, s1) of level l: one chunk of a direction's walk.
struct Chunk {
  int l, s0, s1;
};

// The chunk after c (the rest of its level, or the first chunk of the next
// level; levels are never empty); c itself, with more = false, after the last.
__device__ __forceinline__ Chunk next_chunk(const int32_t* __restrict__ off, int nl, int chunk, const Chunk& c,
                                            bool& more) {
  const bool last = c.s1 == off[c.l + 1];
  const int ln = last ? c.l + 1 : c.l;
  more = ln < nl;
  if (!more) return c;
  const int n0 = last ? off[ln] : c.s1;
  return Chunk{ln, n0, min(n0 + chunk, off[ln + 1])};
}

// One chunk of the two-stage walk.  On entry P holds the current chunk (records
// decoded, cross sections and sources loaded one chunk ago) and raw the records
// of the next chunk (loaded one chunk ago).  The chunk's upstream traces are read
// first, LDS slots before global memory, so that the LDS reads do not wait on
// the vector-memory counter for loads into the same registers.  Then raw is
// decoded into Pn and the next chunk's cross sections and sources are issued, the
// records of the chunk after it go into raw, the staging copy for the next level
// is issued, and only then are the current cells solved: every load a solve
// waits for was issued a chunk earlier, and no load address waits on a load of
// the same chunk.  The look-ahead issues the same loads on every path (lanes past
// a chunk's end re-read its last cell).
template <int R, int LDS, bool NT = false>
__device__ __forceinline__ bool chunk_two_stage(const SweepArgs& A, int i, const int32_t* __restrict__ off, int nl,
                                                int chunk, int cpr, int c, int g, bool lane_ok, Chunk& cur, Chunk& nx,
                                                bool& has_nx, const CellPrep (&P)[R], CellPrep (&Pn)[R],
                                                const RecRaw (&raw)[R], RecRaw (&raw_next)[R], const TraceRing& ring,
                                                const int32_t* __restrict__ so) {
  constexpr int kSolve = LDS ? kTraceGlobalRing : kTraceGlobal;
  double u[R][4][3];
  unsigned hit = 0;
  if constexpr (LDS) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int s = cur.s0 + c + r * cpr;
      if (lane_ok && s < cur.s1 && ring_hit(P[r])) {
        hit |= 1u << r;
        load_upstream<kTraceRing>(A, i, P[r], g, ring, u[r]);
      }
    }
    // opaque to the optimiser: the global reads stay a branch of their own after the LDS reads
    asm volatile("" : "+v"(hit));
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int s = cur.s0 + c + r * cpr;
    if (lane_ok && s < cur.s1 && !(hit >> r & 1u)) load_upstream<kSolve>(A, i, P[r], g, ring, u[r]);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) decode_rec<NT, LDS>(A, i, min(nx.s0 + c + r * cpr, nx.s1 - 1), g, raw[r], Pn[r]);
  bool more2 = false;
  const Chunk nn = has_nx ? next_chunk(off, nl, chunk, nx, more2) : nx;
#pragma unroll
  for (int r = 0; r < R; ++r) load_rec<NT, LDS>(A, i, min(nn.s0 + c + r * cpr, nn.s1 - 1), raw_next[r]);
  const bool last = cur.s1 == off[cur.l + 1];
  const StageCopy sc = stage_load<LDS>(A, i, so, last ? cur.l + 1 : -1);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int s = cur.s0 + c + r * cpr;
    if (lane_ok && s < cur.s1) solve_with_upstream<kSolve, LDS>(A, i, s, P[r], g, ring, u[r]);
  }
  stage_store<LDS>(ring, sc);
  if (last) __syncthreads();
  if (!has_nx) return false;
  cur = nx;
  nx = nn;
  has_nx = more2;
  return true;
}

// Direction walk with a two-stage look-ahead (few groups per cell): records two
// chunks ahead of their solve, cross sections and sources one chunk ahead.
// Measured on 384k tets x 256 directions against a one-stage look-ahead (whose
// source loads waited on the record loads of the same chunk): one group, LDS
// kernel at 512 threads 5.07 -> 4.70 ms per sweep inside the solve, at 256
// threads 6.23 -> 5.62 ms; plain direction kernel at 256 threads 6.79 -> 6.43
// ms; two groups at 512 threads 8.58 -> 8.13 ms.
template <int BLOCK, int R, int LDS = 0, bool NT = false>
__device__ __forceinline__ void walk_levels_two_stage(const SweepArgs& A, int i, const int32_t* __restrict__ off,
                                                      int nl, TraceRing ring = TraceRing{nullptr},
                                                      const int32_t* __restrict__ so = nullptr) {
  const int cpr = BLOCK / A.G;
  const int c = threadIdx.x / A.G;
  const int g = threadIdx.x - c * A.G;
  const bool lane_ok = c < cpr;
  const int chunk = R * cpr;
  if (nl <= 0) return;
  CellPrep Pa[R], Pb[R];
  RecRaw ra[R], rb[R];
  Chunk cur{0, off[0], min(off[0] + chunk, off[1])};
  bool has_nx = false;
  Chunk nx = next_chunk(off, nl, chunk, cur, has_nx);
#pragma unroll
  for (int r = 0; r < R; ++r) prep_cell<NT, LDS>(A, i, min(cur.s0 + c + r * cpr, cur.s1 - 1), g, Pa[r]);
#pragma unroll
  for (int r = 0; r < R; ++r) load_rec<NT, LDS>(A, i, min(nx.s0 + c + r * cpr, nx.s1 - 1), ra[r]);
  while (true) {
    if (!chunk_two_stage<R, LDS, NT>(A, i, off, nl, chunk, cpr, c, g, lane_ok, cur, nx, has_nx, Pa, Pb, ra, rb, ring,
                                     so))
      break;
    if (!chunk_two_stage<R, LDS, NT>(A, i, off, nl, chunk, cpr, c, g, lane_ok, cur, nx, has_nx, Pb, Pa, rb, ra, ring,
                                     so))
      break;
  }
}

// Rounds of static data held per lane for a workgroup size (register budget:
// a 1024-lane workgroup leaves 128 VGPRs per lane; a 768-lane one 168, where one
// round of the two-stage walk fits in 146 and three waves share each SIMD; a
// 256-lane one runs one wave per SIMD and holds four rounds in 256 VGPRs
// without spills).  The LDS sweep of one group at 768 threads against 512
// (two rounds, two waves per SIMD): the 24-group solve 1.241 s against 1.319-
// 1.334 s on one box.
template <int BLOCK>
struct LevelRounds {
  static constexpr int value = BLOCK >= 768 ? 1 : (BLOCK >= 512 ? 2 : 4);
};

// Direction-resident sweep: one workgroup owns one direction (of one batch of
// groups) and walks its levels in order.  All producers and consumers of a
// direction's upstream traces run on one CU, so workgroup-scope ordering is
// enough: no inter-workgroup flags, no agent-scope cache write-back or
// invalidation, and the traces of the last few levels are re-read from the
// XCD's L2.  One launch sweeps every direction.
template <int BLOCK, bool NT>
__global__ __launch_bounds__(BLOCK) void sweep_direction_kernel(const int32_t* __restrict__ lvl_off,
                                                                const int32_t* __restrict__ nlev, int maxlev,
                                                                SweepArgs A) {
  const int i = blockIdx.x;
  // the two-stage walk at any group count (8 groups, the multigroup solve's tail, at 512 threads: 19.6 -> 17.3 ms
  // per sweep of 384k tets x 256 directions against one step per level; 24 groups 57.4 -> 54.8 ms at 512
  // threads, 53.7 -> 55.7 ms at 1024, where level launches stay faster)
  walk_levels_two_stage<BLOCK, LevelRounds<BLOCK>::value, kLdsNone, NT>(A, i, lvl_off + (int64_t)i * (maxlev + 1),
                                                                    nlev[i]);
}

// Direction-resident sweep with LDS trace slots (dynamic shared memory of
// (ring + 2 stage) * 4 * G doubles, the slot plan of this group-batch size):
// upstream traces come from LDS where the plan holds them, otherwise from global
// memory.  The flux is written to both; the workgroup barrier between levels
// orders the LDS writes and reads.
template <int BLOCK, bool NT, bool F32 = false>
__global__ __launch_bounds__(BLOCK) void sweep_direction_lds_kernel(const int32_t* __restrict__ lvl_off,
                                                                    const int32_t* __restrict__ nlev, int maxlev,
                                                                    const int32_t* __restrict__ stage_off,
                                                                    SweepArgs A) {
  extern __shared__ double ring_buf[];
  const int i = blockIdx.x;
  // the staging offsets as a kernel parameter of their own: a restrict pointer the
  // kernel never stores through, so their wave-uniform reads are scalar loads and
  // do not wait on the vector-memory counter, which at each level's end would also
  // wait for the look-ahead loads just issued (a restrict member of SweepArgs does
  // not carry that guarantee to the compiler)
  const int32_t* so = stage_off + (int64_t)i * (maxlev + 2);
  if constexpr (NT)
    walk_levels<BLOCK, LevelRounds<BLOCK>::value, NT, kLdsPlan>(A, i, lvl_off + (int64_t)i * (maxlev + 1), nlev[i],
                                                                TraceRing{ring_buf}, so);
  else
    walk_levels_two_stage<BLOCK, LevelRounds<BLOCK>::value, F32 ? kLdsF32 : kLdsPlan>(
        A, i, lvl_off + (int64_t)i * (maxlev + 1), nlev[i], TraceRing{ring_buf}, so);
}

// Persistent sweep of a batch of independent problems: a fixed grid (a few
// workgroups per CU) pulls (problem, direction) items from a global counter, in
// the order of the item table (largest problems first), and walks each item's
// levels like the direction-resident kernel.  Items of inactive problems
// (converged in a batched solve) are skipped.  Every workgroup leaves the loop
// once the counter passes the item count.
template <int BLOCK, bool NT>
__global__ __launch_bounds__(BLOCK) void sweep_persistent_kernel(const int4* __restrict__ items, int nitems,
                                                                 const int32_t* __restrict__ lvl,
                                                                 const int32_t* __restrict__ active,
                                                                 unsigned int* __restrict__ counter, SweepArgs A) {
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= nitems) break;
    const int4 d = items[it];
    if (active[d.x] == 0) continue;
    // one round: several workgroups share a CU (two-stage walk, few groups: 10.96-11.01 ms per sweep of the
    // 64-problem batch against 10.93-10.94 ms with the one-stage walk it replaced, same box)
    if constexpr (NT)
      walk_levels<BLOCK, 1, NT>(A, d.y, lvl + d.z, d.w);
    else
      walk_levels_two_stage<BLOCK, 1>(A, d.y, lvl + d.z, d.w);
  }
}

__global__ void reset_counter_kernel(unsigned int* counter) {
  if (threadIdx.x == 0) counter[0] = 0u;
}

// inv[i][e] = sweep position of element e in direction i's device order.
__global__ __launch_bounds__(kBlock) void inverse_order_kernel(int NI, int64_t NE, const uint32_t* __restrict__ tasks,
                                                               int32_t* __restrict__ inv) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= (int64_t)NI * NE) return;
  const int64_t i = k / NE;
  inv[i * NE + (tasks[k] & 0xFFFFFFu)] = (int32_t)(k - i * NE);
}

// Staging entries (LDS slot, element) of the LDS sweep's plan with the element
// replaced by its sweep position: one workgroup per direction, whose entries
// are so[0] .. so[stride - 1].
__global__ __launch_bounds__(kBlock) void stage_positions_kernel(int64_t NE, int64_t stride,
                                                                 const int32_t* __restrict__ soff,
                                                                 const int2* __restrict__ stage,
                                                                 const int32_t* __restrict__ inv,
                                                                 int2* __restrict__ out) {
  const int64_t i = blockIdx.x;
  const int32_t* so = soff + i * stride;
  for (int64_t j = so[0] + threadIdx.x; j < so[stride - 1]; j += kBlock) {
    const int2 en = stage[j];
    out[j] = make_int2(en.x, inv[i * NE + en.y]);
  }
}

// One group's flux from sweep order to the natural layout: lane per (direction,
// element, half cell), 16 bytes each.
__global__ __launch_bounds__(kBlock) void flux_to_natural_kernel(int NI, int64_t NE, const int32_t* __restrict__ inv,
                                                                 const double* __restrict__ psi_so,
                                                                 double* __restrict__ psi) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= (int64_t)NI * NE * 2) return;
  const int64_t ie = k >> 1, i = ie / NE;
  const int h = int(k & 1) * 2;
  typedef double d2 __attribute__((ext_vector_type(2)));
  *reinterpret_cast<d2*>(psi + ie * 4 + h) =
      *reinterpret_cast<const d2*>(psi_so + (i * NE + inv[ie]) * 4 + h);
}

// Workgroup index with consecutive blocks on one XCD: the dispatcher deals
// workgroups to the 8 XCDs in turn, so block b runs on XCD b mod 8; this maps
// the workgroups of each XCD to one contiguous range of blocks (the last
// gridDim.x mod 8 blocks keep their index).  Kernels that gather or scatter by
// sweep position use it so that the cells of neighbouring blocks, whose
// positions share cache lines, meet in the same L2.
__device__ __forceinline__ int64_t xcd_block() {
  const unsigned b = blockIdx.x, per = gridDim.x >> 3;
  if (b >= per << 3) return b;
  return (int64_t)(b & 7u) * per + (b >> 3);
}

// phi[lm][col] = sum_i D[lm][i] psi[i][col]; one lane per column (column =
// spatial dof x group), D read through the scalar cache.
template <int NM>
__global__ __launch_bounds__(kBlock) void moments_kernel(const double* __restrict__ D, int NI, int64_t ncols,
                                                         const double* __restrict__ psi, double* __restrict__ phi) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= ncols) return;
  double acc[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) acc[m] = 0.0;
  for (int i = 0; i < NI; ++i) {
    const double v = psi[(int64_t)i * ncols + col];
#pragma unroll
    for (int m = 0; m < NM; ++m) acc[m] += D[m * NI + i] * v;
  }
#pragma unroll
  for (int m = 0; m < NM; ++m) phi[(int64_t)m * ncols + col] = acc[m];
}

// psi[i][col] (+)= sum_lm M[i][lm] sigma_{l(lm)}(mat, g) phi[lm][col].
template <int NM>
__global__ __launch_bounds__(kBlock) void moments_to_discrete_kernel(
    const double* __restrict__ M, const int32_t* __restrict__ deg, int NI, int64_t ncols, int G, int Gtot, int g0,
    int L, const int32_t* __restrict__ mat, const double* __restrict__ sigs, const double* __restrict__ phi,
    double* __restrict__ psi, int accumulate) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= ncols) return;
  const int64_t sd = col / G;
  const int g = int(col - sd * G);
  double sp[NM];
  if (sigs) {
    const int m_el = mat[sd >> 2];
    const double* sl = sigs + ((int64_t)m_el * Gtot + g0 + g) * (L + 1);
#pragma unroll
    for (int m = 0; m < NM; ++m) sp[m] = phi[(int64_t)m * ncols + col] * sl[deg[m]];
  } else {
#pragma unroll
    for (int m = 0; m < NM; ++m) sp[m] = phi[(int64_t)m * ncols + col];
  }
  for (int i = 0; i < NI; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) acc += M[i * NM + m] * sp[m];
    double* dst = psi + (int64_t)i * ncols + col;
    *dst = accumulate ? *dst + acc : acc;
  }
}

// Two-column variants of the two products above for up to 16 moments (P3), the
// solve's hot path: a lane owns columns c, c + 1 and moves them as 16-byte
// pieces; psi is read (D psi) or written (M phi) once per product and nothing
// reuses it from cache, so those accesses are non-temporal.  D psi issues the
// loads of kMomRows directions before their FMAs.  On 384k tets x 256
// directions: D psi 0.59 ms, M phi 0.58 ms, against 0.63 / 0.67 ms for rocBLAS
// dgemm and 0.90 / 0.75 ms for one lane per column.  Each column's sum runs
// over directions (moments) in index order, as in the one-column kernels.
typedef double dbl2 __attribute__((ext_vector_type(2)));
constexpr int kMomRows = 4;

template <int NM>
__global__ __launch_bounds__(kBlock) void moments_pair_kernel(const double* __restrict__ D, int NI, int64_t ncols,
                                                              const double* __restrict__ psi,
                                                              double* __restrict__ phi) {
  const int64_t c0 = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 2;
  if (c0 >= ncols) return;
  double a0[NM], a1[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) a0[m] = a1[m] = 0.0;
  const dbl2* p = reinterpret_cast<const dbl2*>(psi + c0);
  const int64_t stride = ncols / 2;
  int i0 = 0;
  for (; i0 + kMomRows <= NI; i0 += kMomRows) {
    dbl2 v[kMomRows];
#pragma unroll
    for (int r = 0; r < kMomRows; ++r) v[r] = __builtin_nontemporal_load(p + (int64_t)(i0 + r) * stride);
#pragma unroll
    for (int r = 0; r < kMomRows; ++r)
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const double d = D[m * NI + i0 + r];
        a0[m] = fma(d, v[r].x, a0[m]);
        a1[m] = fma(d, v[r].y, a1[m]);
      }
  }
  for (; i0 < NI; ++i0) {
    const dbl2 v = __builtin_nontemporal_load(p + (int64_t)i0 * stride);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const double d = D[m * NI + i0];
      a0[m] = fma(d, v.x, a0[m]);
      a1[m] = fma(d, v.y, a1[m]);
    }
  }
#pragma unroll
  for (int m = 0; m < NM; ++m) *reinterpret_cast<dbl2*>(phi + (int64_t)m * ncols + c0) = dbl2{a0[m], a1[m]};
}

// D psi for one group whose flux is in sweep order: the lane of columns c, c + 1
// (nodes k, k + 1 of cell e) reads them in direction i at e's sweep position
// inv[i][e].  The positions of kGatherRows directions are loaded one batch
// ahead of the flux they address.  Same sums in the same order as
// moments_pair_kernel, so the moments are bit-identical.
constexpr int kGatherRows = 8;

template <int NM>
__global__ __launch_bounds__(kBlock) void moments_gather_pair_kernel(const double* __restrict__ D, int NI, int64_t NE,
                                                                     const int32_t* __restrict__ inv,
                                                                     const double* __restrict__ psi_so,
                                                                     double* __restrict__ phi) {
  const int64_t ncols = NE * 4;
  const int64_t c0 = (xcd_block() * kBlock + threadIdx.x) * 2;
  if (c0 >= ncols) return;
  const int64_t e = c0 >> 2;
  const int h = int(c0 & 3);
  double a0[NM], a1[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) a0[m] = a1[m] = 0.0;
  const int32_t* ip = inv + e;
  const double* ps = psi_so + h;
  int pn[kGatherRows];
  const int nfull = NI - NI % kGatherRows;
  if (nfull > 0) {
#pragma unroll
    for (int r = 0; r < kGatherRows; ++r) pn[r] = __builtin_nontemporal_load(ip + (int64_t)r * NE);
  }
  for (int i0 = 0; i0 < nfull; i0 += kGatherRows) {
    dbl2 v[kGatherRows];
#pragma unroll
    for (int r = 0; r < kGatherRows; ++r)
      v[r] = *reinterpret_cast<const dbl2*>(ps + ((int64_t)(i0 + r) * NE + pn[r]) * 4);
    if (i0 + kGatherRows < nfull) {
#pragma unroll
      for (int r = 0; r < kGatherRows; ++r)
        pn[r] = __builtin_nontemporal_load(ip + (int64_t)(i0 + kGatherRows + r) * NE);
    }
#pragma unroll
    for (int r = 0; r < kGatherRows; ++r)
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const double d = D[m * NI + i0 + r];
        a0[m] = fma(d, v[r].x, a0[m]);
        a1[m] = fma(d, v[r].y, a1[m]);
      }
  }
  for (int i0 = nfull; i0 < NI; ++i0) {
    const dbl2 v = *reinterpret_cast<const dbl2*>(ps + ((int64_t)i0 * NE + ip[(int64_t)i0 * NE]) * 4);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const double d = D[m * NI + i0];
      a0[m] = fma(d, v.x, a0[m]);
      a1[m] = fma(d, v.y, a1[m]);
    }
  }
#pragma unroll
  for (int m = 0; m < NM; ++m) *reinterpret_cast<dbl2*>(phi + (int64_t)m * ncols + c0) = dbl2{a0[m], a1[m]};
}

template <int NM>
__global__ __launch_bounds__(kBlock) void moments_to_discrete_pair_kernel(
    const double* __restrict__ M, const int32_t* __restrict__ deg, int NI, int64_t ncols, int G, int Gtot, int g0,
    int L, const int32_t* __restrict__ mat, const double* __restrict__ sigs, const double* __restrict__ phi,
    double* __restrict__ psi, int accumulate) {
  const int64_t c0 = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 2;
  if (c0 >= ncols) return;
  double s0[NM], s1[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const dbl2 v = *reinterpret_cast<const dbl2*>(phi + (int64_t)m * ncols + c0);
    s0[m] = v.x;
    s1[m] = v.y;
  }
  if (sigs) {
    const int64_t sd0 = c0 / G, sd1 = (c0 + 1) / G;
    const int ga = int(c0 - sd0 * G), gb = int(c0 + 1 - sd1 * G);
    const double* la = sigs + ((int64_t)mat[sd0 >> 2] * Gtot + g0 + ga) * (L + 1);
    const double* lb = sigs + ((int64_t)mat[sd1 >> 2] * Gtot + g0 + gb) * (L + 1);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      s0[m] *= la[deg[m]];
      s1[m] *= lb[deg[m]];
    }
  }
  dbl2* o = reinterpret_cast<dbl2*>(psi + c0);
  const int64_t stride = ncols / 2;
  for (int i = 0; i < NI; ++i) {
    double x0 = 0.0, x1 = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const double mm = M[i * NM + m];
      x0 = fma(mm, s0[m], x0);
      x1 = fma(mm, s1[m], x1);
    }
    dbl2* dst = o + (int64_t)i * stride;
    if (accumulate) {
      const dbl2 old = *dst;
      x0 = old.x + x0;
      x1 = old.y + x1;
    }
    __builtin_nontemporal_store(dbl2{x0, x1}, dst);
  }
}

// D psi for one group whose flux is FP32 in sweep order (the FP32 flux mode):
// as moments_gather_pair_kernel, 8-byte pieces converted to FP64 before the
// FP64 sums.
template <int NM>
__global__ __launch_bounds__(kBlock) void moments_gather32_pair_kernel(const double* __restrict__ D, int NI,
                                                                       int64_t NE, const int32_t* __restrict__ inv,
                                                                       const float* __restrict__ psi32,
                                                                       double* __restrict__ phi) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  const int64_t ncols = NE * 4;
  const int64_t c0 = (xcd_block() * kBlock + threadIdx.x) * 2;
  if (c0 >= ncols) return;
  const int64_t e = c0 >> 2;
  const int h = int(c0 & 3);
  double a0[NM], a1[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) a0[m] = a1[m] = 0.0;
  const int32_t* ip = inv + e;
  const float* ps = psi32 + h;
  int pn[kGatherRows];
  const int nfull = NI - NI % kGatherRows;
  if (nfull > 0) {
#pragma unroll
    for (int r = 0; r < kGatherRows; ++r) pn[r] = __builtin_nontemporal_load(ip + (int64_t)r * NE);
  }
  for (int i0 = 0; i0 < nfull; i0 += kGatherRows) {
    f32x2 v[kGatherRows];
#pragma unroll
    for (int r = 0; r < kGatherRows; ++r)
      v[r] = *reinterpret_cast<const f32x2*>(ps + ((int64_t)(i0 + r) * NE + pn[r]) * 4);
    if (i0 + kGatherRows < nfull) {
#pragma unroll
      for (int r = 0; r < kGatherRows; ++r)
        pn[r] = __builtin_nontemporal_load(ip + (int64_t)(i0 + kGatherRows + r) * NE);
    }
#pragma unroll
    for (int r = 0; r < kGatherRows; ++r) {
      const double x0 = (double)v[r].x, x1 = (double)v[r].y;
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const double d = D[m * NI + i0 + r];
        a0[m] = fma(d, x0, a0[m]);
        a1[m] = fma(d, x1, a1[m]);
      }
    }
  }
  for (int i0 = nfull; i0 < NI; ++i0) {
    const f32x2 v = *reinterpret_cast<const f32x2*>(ps + ((int64_t)i0 * NE + ip[(int64_t)i0 * NE]) * 4);
    const double x0 = (double)v.x, x1 = (double)v.y;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const double d = D[m * NI + i0];
      a0[m] = fma(d, x0, a0[m]);
      a1[m] = fma(d, x1, a1[m]);
    }
  }
#pragma unroll
  for (int m = 0; m < NM; ++m) *reinterpret_cast<dbl2*>(phi + (int64_t)m * ncols + c0) = dbl2{a0[m], a1[m]};
}

// M S phi for one group written in sweep order: psi_so[i][inv[i][e]] gets cell
// e's values of direction i (the same sums as moments_to_discrete_pair_kernel,
// so bit-identical values, at other addresses).  Neighbouring cells of one
// level sit at consecutive positions; with consecutive blocks on one XCD their
// 32-byte pieces meet in its L2 before they are written back.
template <int NM>
__global__ __launch_bounds__(kBlock) void moments_to_order_pair_kernel(
    const double* __restrict__ M, const int32_t* __restrict__ deg, int NI, int64_t NE, int Gtot, int g0, int L,
    const int32_t* __restrict__ mat, const double* __restrict__ sigs, const int32_t* __restrict__ inv,
    const double* __restrict__ phi, double* __restrict__ psi_so) {
  const int64_t ncols = NE * 4;
  const int64_t c0 = (xcd_block() * kBlock + threadIdx.x) * 2;
  if (c0 >= ncols) return;
  const int64_t e = c0 >> 2;
  const int h = int(c0 & 3);
  double s0[NM], s1[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const dbl2 v = *reinterpret_cast<const dbl2*>(phi + (int64_t)m * ncols + c0);
    s0[m] = v.x;
    s1[m] = v.y;
  }
  if (sigs) {
    const double* la = sigs + ((int64_t)mat[e] * Gtot + g0) * (L + 1);
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      s0[m] *= la[deg[m]];
      s1[m] *= la[deg[m]];
    }
  }
  const int32_t* ip = inv + e;
  for (int i = 0; i < NI; ++i) {
    double x0 = 0.0, x1 = 0.0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      const double mm = M[i * NM + m];
      x0 = fma(mm, s0[m], x0);
      x1 = fma(mm, s1[m], x1);
    }
    const int p = __builtin_nontemporal_load(ip + (int64_t)i * NE);
    *reinterpret_cast<dbl2*>(psi_so + ((int64_t)i * NE + p) * 4 + h) = dbl2{x0, x1};
  }
}

// Scattering orders above 7 ((L+1)^2 > 64 moments) do not fit one lane's
// accumulators: the moments are split into chunks of kMomChunk, one grid row
// of workgroups per chunk for D psi, and one accumulating pass per chunk for
// M phi.
constexpr int kMomChunk = 32;

// phi[m0 + m][col] = sum_i D[m0 + m][i] psi[i][col], m < kMomChunk, m0 = blockIdx.y kMomChunk.
__global__ __launch_bounds__(kBlock) void moments_chunk_kernel(const double* __restrict__ D, int NI, int nmom,
                                                               int64_t ncols, const double* __restrict__ psi,
                                                               double* __restrict__ phi) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= ncols) return;
  const int m0 = (int)blockIdx.y * kMomChunk;
  double acc[kMomChunk];
#pragma unroll
  for (int m = 0; m < kMomChunk; ++m) acc[m] = 0.0;
  for (int i = 0; i < NI; ++i) {
    const double v = psi[(int64_t)i * ncols + col];
#pragma unroll
    for (int m = 0; m < kMomChunk; ++m) acc[m] += D[(int64_t)min(m0 + m, nmom - 1) * NI + i] * v;
  }
#pragma unroll
  for (int m = 0; m < kMomChunk; ++m)
    if (m0 + m < nmom) phi[(int64_t)(m0 + m) * ncols + col] = acc[m];
}

// psi[i][col] (+)= sum_{m0 <= lm < m0 + kMomChunk} M[i][lm] sigma_{l(lm)}(mat, g) phi[lm][col].
__global__ __launch_bounds__(kBlock) void moments_to_discrete_chunk_kernel(
    const double* __restrict__ M, const int32_t* __restrict__ deg, int NI, int nmom, int m0, int64_t ncols, int G,
    int Gtot, int g0, int L, const int32_t* __restrict__ mat, const double* __restrict__ sigs,
    const double* __restrict__ phi, double* __restrict__ psi, int accumulate) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= ncols) return;
  const int64_t sd = col / G;
  const int g = int(col - sd * G);
  const double* sl = sigs ? sigs + ((int64_t)mat[sd >> 2] * Gtot + g0 + g) * (L + 1) : nullptr;
  double sp[kMomChunk];
#pragma unroll
  for (int m = 0; m < kMomChunk; ++m) {
    const int lm = m0 + m;
    sp[m] = lm < nmom ? phi[(int64_t)lm * ncols + col] * (sl ? sl[deg[lm]] : 1.0) : 0.0;
  }
  for (int i = 0; i < NI; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < kMomChunk; ++m) acc += M[(int64_t)i * nmom + min(m0 + m, nmom - 1)] * sp[m];
    double* dst = psi + (int64_t)i * ncols + col;
    *dst = accumulate ? *dst + acc : acc;
  }
}

// out[lm][col] = sigma_{l(lm)}(mat, g) phi[lm][col]: the input of the M GEMM.
__global__ __launch_bounds__(kBlock) void scale_moments_kernel(const double* __restrict__ phi, int nmom,
                                                               int64_t ncols, int G, int Gtot, int g0, int L,
                                                               const int32_t* __restrict__ deg,
                                                               const int32_t* __restrict__ mat,
                                                               const double* __restrict__ sigs,
                                                               double* __restrict__ out) {
  const int64_t n = (int64_t)nmom * ncols;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
    const int lm = int(j / ncols);
    const int64_t col = j - (int64_t)lm * ncols;
    const int64_t sd = col / G;
    const int g = int(col - sd * G);
    out[j] = phi[j] * sigs[((int64_t)mat[sd >> 2] * Gtot + g0 + g) * (L + 1) + deg[lm]];
  }
}

// Dense kernel of ScatteringAlgebra.f90:30-53: y[i][col] = sum_ip x[ip][col] YYS_m[ip][i],
// m = material of the column; 16 outputs per pass over the column.
__global__ __launch_bounds__(kBlock) void scatter_dense_kernel(const double* __restrict__ yys, int NI, int64_t ncols,
                                                               int G, const int32_t* __restrict__ mat,
                                                               const double* __restrict__ x, double* __restrict__ y) {
  const int64_t col = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (col >= ncols) return;
  const int m_el = mat[(col / G) >> 2];
  const double* Y = yys + (int64_t)m_el * NI * NI;
  for (int i0 = 0; i0 < NI; i0 += 16) {
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0;
    const int nj = min(16, NI - i0);
    for (int ip = 0; ip < NI; ++ip) {
      const double v = x[(int64_t)ip * ncols + col];
      const double* row = Y + (int64_t)ip * NI + i0;
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j < nj) acc[j] += v * row[j];
    }
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (j < nj) y[(int64_t)(i0 + j) * ncols + col] = acc[j];
  }
}

// Energy-coupling source of group g in moment form (the EI loop of
// Main.f90:584-629 with K_{g'->g} = M diag(sigma_l(g'->g)) D):
// acc[lm][sd] = qmom[lm][sd] + sum_{g' < g} sigma_{l(lm)}(g' -> g, mat) phi_{g'}[lm][sd].
__global__ __launch_bounds__(kBlock) void downscatter_kernel(const double* __restrict__ phi_all,
                                                             const double* __restrict__ S, int g, int G, int Nm,
                                                             int nmom, int L, const int32_t* __restrict__ mat,
                                                             const int32_t* __restrict__ deg, int64_t NENK,
                                                             const double* __restrict__ qmom,
                                                             double* __restrict__ acc) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= (int64_t)nmom * NENK) return;
  const int lm = int(j / NENK);
  const int64_t sd = j - (int64_t)lm * NENK;
  const int m_el = mat[sd >> 2];
  const int l = deg[lm];
  double a = qmom ? qmom[j] : 0.0;
  for (int gp = 0; gp < g; ++gp)
    a += S[(((int64_t)gp * G + g) * Nm + m_el) * (L + 1) + l] * phi_all[(int64_t)gp * nmom * NENK + j];
  acc[j] = a;
}

__global__ __launch_bounds__(kBlock) void axpby_kernel(int64_t n, double a, const double* x, double b, const double* y,
                                                       double* z) {
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock)
    z[j] = a * x[j] + b * y[j];
}

__global__ __launch_bounds__(kBlock) void scale_kernel(int64_t n, double a, double* x) {
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) x[j] *= a;
}

// Deterministic two-stage dot product: fixed grid, fixed reduction tree.
constexpr int kDotBlocks = kDotPartials;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

__global__ __launch_bounds__(kBlock) void dot_partial_kernel(int64_t n, const double* __restrict__ x,
                                                             const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock)
    acc += x[j] * y[j];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void dot_final_kernel(int nparts, const double* __restrict__ part,
                                                           double* __restrict__ out) {
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int j = threadIdx.x; j < nparts; j += kBlock) acc += part[j];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = s;
}

// Fused dots and updates against up to kOrthoChunk basis vectors at a time
// (V_i = V + i ld): one read of w serves every dot of a chunk, and one
// read-write of w every subtraction.  Partial sums per block reduce inside each
// wave by a fixed butterfly and across the block's waves in wave order, then
// over blocks in block order, so results do not vary from run to run.
constexpr int kOrthoChunk = 8;
constexpr int kOrthoBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// A kernel given a gate (state word, mask) runs only when bit *state of mask is
// set: the device-side Arnoldi step below switches its optional second
// Gram-Schmidt pass on and off this way, without a host round trip.
__device__ __forceinline__ bool gate_closed(const int* state, int mask) {
  return state != nullptr && ((mask >> *state) & 1) == 0;
}

// part[(slot0 + i) kOrthoBlocks + block] = partial <w, V_i> for i < c; with
// self_slot >= 0 also part[self_slot kOrthoBlocks + block] = partial <w, w>.
__global__ __launch_bounds__(kBlock) void multi_dot_partial_kernel(int64_t n, int c, const double* __restrict__ V,
                                                                   int64_t ld, const double* __restrict__ w,
                                                                   int slot0, int self_slot,
                                                                   double* __restrict__ part,
                                                                   const int* gate, int mask) {
  if (gate_closed(gate, mask)) return;
  __shared__ double sh[kOrthoChunk + 1][kBlock / 64];
  double acc[kOrthoChunk + 1];
#pragma unroll
  for (int i = 0; i <= kOrthoChunk; ++i) acc[i] = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
    const double wj = w[j];
#pragma unroll
    for (int i = 0; i < kOrthoChunk; ++i)
      if (i < c) acc[i] += wj * V[i * ld + j];
    acc[kOrthoChunk] += wj * wj;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i <= kOrthoChunk; ++i) {
    const double s = wave_sum(acc[i]);
    if (lane == 0) sh[i][wave] = s;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < c || (t == c && self_slot >= 0)) {
    const int i = t < c ? t : kOrthoChunk;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) s += sh[i][q];
    part[(int64_t)(t < c ? slot0 + t : self_slot) * kOrthoBlocks + blockIdx.x] = s;
  }
}

// out[slot] = sum over blocks of part[slot kOrthoBlocks + block]; one block per slot.
__global__ __launch_bounds__(kBlock) void multi_dot_final_kernel(const double* __restrict__ part,
                                                                 double* __restrict__ out, const int* gate,
                                                                 int mask) {
  if (gate_closed(gate, mask)) return;
  __shared__ double sh[kBlock];
  double acc = 0.0;
  for (int j = threadIdx.x; j < kOrthoBlocks; j += kBlock) acc += part[(int64_t)blockIdx.x * kOrthoBlocks + j];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

struct OrthoCoef {
  double h[kOrthoChunk];
};

// w <- (w - sum_{i<c} h_i V_i) * scale
__global__ __launch_bounds__(kBlock) void multi_axpy_kernel(int64_t n, int c, const double* __restrict__ V, int64_t ld,
                                                            OrthoCoef h, double scale, double* __restrict__ w) {
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
    double a = w[j];
#pragma unroll
    for (int i = 0; i < kOrthoChunk; ++i)
      if (i < c) a -= h.h[i] * V[i * ld + j];
    w[j] = a * scale;
  }
}

// The same update with the coefficients (and the scale, when given) in device
// memory, written there by the kernels below.
__global__ __launch_bounds__(kBlock) void multi_axpy_dev_kernel(int64_t n, int c, const double* __restrict__ V,
                                                                int64_t ld, const double* __restrict__ coef,
                                                                const double* __restrict__ scale_p,
                                                                const int* gate, int mask, double* __restrict__ w) {
  if (gate_closed(gate, mask)) return;
  double h[kOrthoChunk];
#pragma unroll
  for (int i = 0; i < kOrthoChunk; ++i) h[i] = i < c ? coef[i] : 0.0;
  const double scale = scale_p ? *scale_p : 1.0;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
    double a = w[j];
#pragma unroll
    for (int i = 0; i < kOrthoChunk; ++i)
      if (i < c) a -= h[i] * V[i * ld + j];
    w[j] = a * scale;
  }
}

// The update of multi_axpy_dev_kernel for at most kOrthoChunk vectors fused with
// the dots of the next Gram-Schmidt pass (when the state asks for one): each
// element of the updated w meets the V_i it was just updated with, so the pass
// reads V and w once instead of twice.  Runs on the fixed grid of
// multi_dot_partial_kernel (kOrthoBlocks) with the same element order per
// thread, so the partial sums, and with them every bit of the dots, are those
// of the separate kernels.
__global__ __launch_bounds__(kBlock) void multi_axpy_dot_kernel(int64_t n, int c, const double* __restrict__ V,
                                                                int64_t ld, const double* __restrict__ coef,
                                                                const double* __restrict__ scale_p, const int* state,
                                                                int mask_axpy, int mask_dot, double* __restrict__ w,
                                                                double* __restrict__ part) {
  if (gate_closed(state, mask_axpy)) return;
  const bool dots = !gate_closed(state, mask_dot);
  __shared__ double sh[kOrthoChunk + 1][kBlock / 64];
  double h[kOrthoChunk], acc[kOrthoChunk + 1];
#pragma unroll
  for (int i = 0; i < kOrthoChunk; ++i) h[i] = i < c ? coef[i] : 0.0;
#pragma unroll
  for (int i = 0; i <= kOrthoChunk; ++i) acc[i] = 0.0;
  const double scale = *scale_p;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock) {
    double v[kOrthoChunk];
    double a = w[j];
#pragma unroll
    for (int i = 0; i < kOrthoChunk; ++i) {
      v[i] = i < c ? V[i * ld + j] : 0.0;
      if (i < c) a -= h[i] * v[i];
    }
    a *= scale;
    w[j] = a;
    if (dots) {
#pragma unroll
      for (int i = 0; i < kOrthoChunk; ++i)
        if (i < c) acc[i] += a * v[i];
      acc[kOrthoChunk] += a * a;
    }
  }
  if (!dots) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i <= kOrthoChunk; ++i) {
    const double s = wave_sum(acc[i]);
    if (lane == 0) sh[i][wave] = s;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t <= c) {
    const int i = t < c ? t : kOrthoChunk;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) s += sh[i][q];
    part[(int64_t)t * kOrthoBlocks + blockIdx.x] = s;
  }
}

// One wave, one lane: the scalar work of an Arnoldi step is a few hundred flops
// in a fixed order, done serially so that its bits never vary.

// After the dots d[0..c) = <w, V_i>, d[c] = <w, w> of Gram-Schmidt pass `pass`:
// accumulate the projections in hk, stage them as the update's coefficients and
// decide, by the test of orthonormalize() (more than half of ||w||^2 left), what
// comes next.  Pass 1 runs only when pass 0 asked for it.
__global__ void ortho_decide_kernel(const double* __restrict__ d, int c, int pass, double* __restrict__ hk,
                                    double* __restrict__ coef, ArnoldiCtl* __restrict__ ctl) {
  if (threadIdx.x != 0) return;
  if (pass == 1 && ctl->state != kArnPass2) return;
  double hh = 0.0;
  for (int i = 0; i < c; ++i) hh += d[i] * d[i];
  const double ww = d[c];
  if (!isfinite(ww) || !isfinite(hh)) {
    ctl->nu = __builtin_nan("");
    ctl->scale = 1.0;
    ctl->state = kArnNaN;
    return;
  }
  for (int i = 0; i < c; ++i) {
    hk[i] = pass == 0 ? d[i] : hk[i] + d[i];
    coef[i] = d[i];
  }
  const double rem = ww - hh;
  if (rem > 0.5 * ww) {
    ctl->nu = sqrt(rem);
    ctl->scale = 1.0 / ctl->nu;
    ctl->state = pass == 0 ? kArnDone1 : kArnDone2;
  } else if (pass == 0) {
    ctl->scale = 1.0;
    ctl->state = kArnPass2;
  } else {
    ctl->nu = 0.0;                        // w lies in span V to working precision
    ctl->state = kArnInvariant;
  }
}

// Givens rotation of Golub & van Loan as SPARSKIT's givens (iters.f:3225-3260).
__device__ void givens_dev(double x, double y, double* c, double* s) {
  if (x == 0.0 && y == 0.0) {
    *c = 1.0;
    *s = 0.0;
  } else if (fabs(y) > fabs(x)) {
    const double t = x / y;
    *s = copysign(1.0 / sqrt(1.0 + t * t), y);
    *c = t * *s;
  } else {
    const double t = y / x;
    *c = copysign(1.0 / sqrt(1.0 + t * t), x);
    *s = t * *c;
  }
}

// Column k of the Hessenberg matrix from the projections hk and the norm nu
// (shifted: of I - K when the basis was built on K), the earlier rotations
// applied to it, its own rotation, and the rotated right-hand side g, whose
// entry k + 1 is the residual estimate (iters.f:2059-2094).
__global__ void arnoldi_givens_kernel(double* __restrict__ H, double* __restrict__ cs, double* __restrict__ sn,
                                      double* __restrict__ g, const double* __restrict__ hk, int k, int ldh,
                                      int shifted, ArnoldiCtl* __restrict__ ctl) {
  if (threadIdx.x != 0) return;
  if (ctl->state == kArnNaN) {
    ctl->resn = __builtin_nan("");
    return;
  }
  const double nu = ctl->nu;
  double* col = H + (int64_t)k * ldh;
  for (int i = 0; i <= k; ++i) col[i] = shifted ? (i == k ? 1.0 : 0.0) - hk[i] : hk[i];
  col[k + 1] = shifted ? -nu : nu;
  for (int i = 0; i < k; ++i) {
    const double a = col[i], b = col[i + 1];
    col[i] = cs[i] * a + sn[i] * b;
    col[i + 1] = cs[i] * b - sn[i] * a;
  }
  givens_dev(col[k], col[k + 1], &cs[k], &sn[k]);
  col[k] = cs[k] * col[k] + sn[k] * col[k + 1];
  col[k + 1] = 0.0;
  g[k + 1] = -sn[k] * g[k];
  g[k] = cs[k] * g[k];
  ctl->resn = fabs(g[k + 1]);
}

__global__ void arnoldi_start_kernel(double* __restrict__ g, int m, double beta) {
  for (int i = threadIdx.x; i <= m; i += blockDim.x) g[i] = i == 0 ? beta : 0.0;
}

// y = -R^-1 g over the first k columns (back substitution), the coefficients of
// the update x <- x - (-V y).
__global__ void arnoldi_solve_kernel(const double* __restrict__ H, const double* __restrict__ g, int k, int ldh,
                                     double* __restrict__ y) {
  if (threadIdx.x != 0) return;
  for (int i = k - 1; i >= 0; --i) {
    double a = g[i];
    for (int j = i + 1; j < k; ++j) a -= H[(int64_t)j * ldh + i] * y[j];
    y[i] = H[(int64_t)i * ldh + i] != 0.0 ? a / H[(int64_t)i * ldh + i] : 0.0;
  }
  for (int i = 0; i < k; ++i) y[i] = -y[i];
}

__global__ __launch_bounds__(kBlock) void fluence_kernel(const double* __restrict__ w, int NI, int64_t NENK,
                                                         const double* __restrict__ v, double* __restrict__ phi) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= NENK) return;
  double acc = 0.0;
  for (int i = 0; i < NI; ++i) acc += w[i] * v[(int64_t)i * NENK + j];
  phi[j] = acc;
}

// x[sd + i NENK] *= xsdd[mat(sd)]: the delta-down kernel, identity in angle
// (ApplyDeltaDown, ScatteringAlgebra.f90:62-90).
__global__ __launch_bounds__(kBlock) void delta_down_kernel(int64_t n, int64_t NENK, const int32_t* __restrict__ mat,
                                                            const double* __restrict__ xsdd, double* __restrict__ x) {
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock)
    x[j] *= xsdd[mat[(j % NENK) >> 2]];
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline unsigned grid_stride(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  return (unsigned)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

}  // namespace

template <bool NT>
static void launch_direction_nt(int block, unsigned ndir, hipStream_t st, const int32_t* lo, const int32_t* nl,
                                int maxlev, const SweepArgs& A) {
  if (block == 256)
    hipLaunchKernelGGL((sweep_direction_kernel<256, NT>), dim3(ndir), dim3(256), 0, st, lo, nl, maxlev, A);
  else if (block == 512)
    hipLaunchKernelGGL((sweep_direction_kernel<512, NT>), dim3(ndir), dim3(512), 0, st, lo, nl, maxlev, A);
  else if (block == 768)
    hipLaunchKernelGGL((sweep_direction_kernel<768, NT>), dim3(ndir), dim3(768), 0, st, lo, nl, maxlev, A);
  else
    hipLaunchKernelGGL((sweep_direction_kernel<1024, NT>), dim3(ndir), dim3(1024), 0, st, lo, nl, maxlev, A);
}

static void launch_direction(int block, bool nt, unsigned ndir, hipStream_t st, const int32_t* lo, const int32_t* nl,
                             int maxlev, const SweepArgs& A) {
  if (nt)
    launch_direction_nt<true>(block, ndir, st, lo, nl, maxlev, A);
  else
    launch_direction_nt<false>(block, ndir, st, lo, nl, maxlev, A);
}

template <bool NT, bool F32 = false>
static int launch_direction_lds_nt(int block, unsigned ndir, hipStream_t st, const int32_t* lo, const int32_t* nl,
                                   int maxlev, const int32_t* so, const SweepArgs& A, size_t bytes) {
  const void* fn = block == 256   ? (const void*)&sweep_direction_lds_kernel<256, NT, F32>
                   : block == 512 ? (const void*)&sweep_direction_lds_kernel<512, NT, F32>
                   : block == 768 ? (const void*)&sweep_direction_lds_kernel<768, NT, F32>
                                  : (const void*)&sweep_direction_lds_kernel<1024, NT, F32>;
  LB_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  if (block == 256)
    hipLaunchKernelGGL((sweep_direction_lds_kernel<256, NT, F32>), dim3(ndir), dim3(256), bytes, st, lo, nl, maxlev,
                       so, A);
  else if (block == 512)
    hipLaunchKernelGGL((sweep_direction_lds_kernel<512, NT, F32>), dim3(ndir), dim3(512), bytes, st, lo, nl, maxlev,
                       so, A);
  else if (block == 768)
    hipLaunchKernelGGL((sweep_direction_lds_kernel<768, NT, F32>), dim3(ndir), dim3(768), bytes, st, lo, nl, maxlev,
                       so, A);
  else
    hipLaunchKernelGGL((sweep_direction_lds_kernel<1024, NT, F32>), dim3(ndir), dim3(1024), bytes, st, lo, nl, maxlev,
                       so, A);
  return LB_OK;
}

static int one_group_vectors(int G, const double* psi) {
  return G == 1 && (reinterpret_cast<uintptr_t>(psi) & 15) == 0;
}

// Slot plan of the LDS sweep for G groups per cell, built on first use from the
// device sweep order (lb_trace_plan.cpp) and kept per group-batch size.  Every
// slot index is checked against the LDS the kernel allocates before upload.
static TracePlan* trace_plan(lb_op* op, int G, int* rc) {
  auto it = op->plans.find(G);
  if (it != op->plans.end()) return &it->second;
  *rc = LB_OK;
  if (op->lds_max == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                                                 hipSuccess) {
      set_error("sweep: cannot query the LDS size");
      *rc = LB_E_HIP;
      return nullptr;
    }
    op->lds_max = v;
  }
  TracePlan p;
  if ((*rc = trace_plan_sizes(op->lds_max, G, &p.ring, &p.stage)) != LB_OK) return nullptr;
  const int64_t NI = op->NI, NE = op->NE, stride = op->maxlev + 2;
  std::vector<uint32_t> tasks(NI * NE);
  if (hipMemcpy(tasks.data(), op->d_tasks.p, tasks.size() * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("sweep: cannot read the device sweep order");
    *rc = LB_E_HIP;
    return nullptr;
  }
  std::vector<int32_t> uslot(NI * NE * 4), wslot(NI * NE), soff(NI * stride), ent;
  std::vector<uint8_t> gprod(G == 1 ? NI * NE : 0);   // one group: the FP32 flux mode's global producers
  if ((*rc = build_trace_plan(NI, NE, op->maxlev, op->h_nlev.data(), op->h_lvl_off.data(), tasks.data(),
                              op->h_nbr.data(), p.ring, p.stage, uslot.data(), wslot.data(), soff.data(), ent,
                              p.stats, G == 1 ? gprod.data() : nullptr)) != LB_OK)
    return nullptr;
  const int slots = p.ring + 2 * p.stage;
  bool ok = slots <= 32767;   // slots travel as 16-bit indices
  for (int32_t v : uslot) ok &= v >= -1 && v < slots;
  for (int32_t v : wslot) ok &= v >= -1 && v < p.ring;
  for (size_t k = 0; k < ent.size(); k += 2) ok &= ent[k] >= p.ring && ent[k] < slots && ent[k + 1] >= 0 && ent[k + 1] < NE;
  for (int64_t i = 0; i < NI; ++i)
    for (int64_t l = 0; l + 1 < stride; ++l) {
      const int32_t n = soff[i * stride + l + 1] - soff[i * stride + l];
      ok &= n >= 0 && n <= p.stage;
    }
  if (!ok) {
    set_error("sweep: LDS slot plan out of range");
    *rc = LB_E_ARG;
    return nullptr;
  }
  if (ent.empty()) ent.assign(2, 0);
  struct Guard {
    TracePlan* t;
    ~Guard() {
      if (t)
        for (DevBuf* b : {&t->d_uslot, &t->d_wslot, &t->d_stage_off, &t->d_stage, &t->d_gprod}) b->release();
    }
  } guard{&p};
  const std::vector<int16_t> us16(uslot.begin(), uslot.end()), ws16(wslot.begin(), wslot.end());
  if ((*rc = upload(op, p.d_uslot, us16.data(), us16.size() * sizeof(int16_t))) != LB_OK) return nullptr;
  if ((*rc = upload(op, p.d_wslot, ws16.data(), ws16.size() * sizeof(int16_t))) != LB_OK) return nullptr;
  if ((*rc = upload(op, p.d_stage_off, soff.data(), soff.size() * sizeof(int32_t))) != LB_OK) return nullptr;
  if ((*rc = upload(op, p.d_stage, ent.data(), ent.size() * sizeof(int32_t))) != LB_OK) return nullptr;
  if (G == 1 && (*rc = upload(op, p.d_gprod, gprod.data(), gprod.size())) != LB_OK) return nullptr;
  guard.t = nullptr;
  return &op->plans.emplace(G, p).first->second;
}

template <bool NT>
static void launch_persistent_nt(int block, dim3 grid, hipStream_t st, const int4* items, int nitems,
                                 const int32_t* lvl, const int32_t* act, unsigned int* ctr, const SweepArgs& A) {
  if (block == 256)
    hipLaunchKernelGGL((sweep_persistent_kernel<256, NT>), grid, dim3(256), 0, st, items, nitems, lvl, act, ctr, A);
  else if (block == 512)
    hipLaunchKernelGGL((sweep_persistent_kernel<512, NT>), grid, dim3(512), 0, st, items, nitems, lvl, act, ctr, A);
  else
    hipLaunchKernelGGL((sweep_persistent_kernel<1024, NT>), grid, dim3(1024), 0, st, items, nitems, lvl, act, ctr, A);
}

static void launch_persistent(int block, bool nt, dim3 grid, hipStream_t st, const int4* items, int nitems,
                              const int32_t* lvl, const int32_t* act, unsigned int* ctr, const SweepArgs& A) {
  if (nt)
    launch_persistent_nt<true>(block, grid, st, items, nitems, lvl, act, ctr, A);
  else
    launch_persistent_nt<false>(block, grid, st, items, nitems, lvl, act, ctr, A);
}

// The records of every (direction, sweep position) for a task order; with inv,
// neighbours by sweep position.
static int ensure_records(lb_op* op, DevBuf& rec, const DevBuf& tasks, const int32_t* inv = nullptr) {
  if (rec.p) return LB_OK;
  const int64_t nrec = op->NI * op->NE;
  LB_TRY(rec.reserve(nrec * sizeof(TaskRec)));
  hipLaunchKernelGGL(build_records_kernel, dim3(grid_for(nrec)), dim3(kBlock), 0, op->stream, (int)op->NI, op->NE,
                     tasks.as<const uint32_t>(), op->d_geo.as<const double4>(), op->d_nbr.as<const int4>(),
                     op->d_perm.as<const uint32_t>(), op->d_mat.as<const int32_t>(), op->d_k.as<const double>(), inv,
                     rec.as<TaskRec>());
  LB_HIP(hipGetLastError());
  return LB_OK;
}

// The sweep position of every element per direction (built on first use): in
// the operator's device order, or with a problem set in the persistent
// sweep's problem-major order.
static int ensure_inverse(lb_op* op) {
  const bool batched = op->batch.P > 0;
  DevBuf& inv = batched ? op->batch.d_inv : op->d_inv;
  if (inv.p) return LB_OK;
  const int64_t n = op->NI * op->NE;
  LB_TRY(inv.reserve(n * sizeof(int32_t)));
  hipLaunchKernelGGL(inverse_order_kernel, dim3(grid_for(n)), dim3(kBlock), 0, op->stream, (int)op->NI, op->NE,
                     (batched ? op->batch.d_tasks : op->d_tasks).as<const uint32_t>(), inv.as<int32_t>());
  LB_HIP(hipGetLastError());
  return LB_OK;
}

static const int32_t* order_inverse(const lb_op* op) {
  return (op->batch.P > 0 ? op->batch.d_inv : op->d_inv).as<const int32_t>();
}

// The staging entries of a slot plan by sweep position (built on first use).
static int ensure_stage_positions(lb_op* op, TracePlan* plan) {
  if (plan->d_stage_pos.p) return LB_OK;
  LB_TRY(ensure_inverse(op));
  LB_TRY(plan->d_stage_pos.reserve(plan->d_stage.bytes));
  hipLaunchKernelGGL(stage_positions_kernel, dim3((unsigned)op->NI), dim3(kBlock), 0, op->stream, op->NE,
                     (int64_t)op->maxlev + 2, plan->d_stage_off.as<const int32_t>(), plan->d_stage.as<const int2>(),
                     order_inverse(op), plan->d_stage_pos.as<int2>());
  LB_HIP(hipGetLastError());
  return LB_OK;
}

// Replays the graph captured for this (source, flux, group range, launch
// configuration), capturing enqueue() first if there is none.  The cache holds 8
// graphs and is cleared whenever the physics table or the problem set moves.
template <class F>
static int replay_graph(lb_op* op, const double* q, double* psi, int g0, int G, const SweepConfig& cfg, int order,
                        F enqueue) {
  hipGraphExec_t exec = nullptr;
  for (auto& sg : op->graphs)
    if (sg.q == q && sg.psi == psi && sg.g0 == g0 && sg.G == G && sg.mode == cfg.mode && sg.block == cfg.block &&
        sg.per_cu == cfg.per_cu && sg.order == order)
      exec = sg.exec;
  if (!exec) {
    LB_HIP(hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal));
    enqueue();
    hipGraph_t graph = nullptr;
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipStreamEndCapture(op->stream, &graph);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      set_error(std::string("sweep graph capture failed: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
      return LB_E_HIP;
    }
    hipError_t e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e3 != hipSuccess) {
      set_error(std::string("sweep graph instantiation failed: ") + hipGetErrorString(e3));
      return LB_E_HIP;
    }
    if (op->graphs.size() >= 8) {
      LB_HIP(hipStreamSynchronize(op->stream));   // the evicted graph may still be in flight
      (void)hipGraphExecDestroy(op->graphs.front().exec);
      op->graphs.erase(op->graphs.begin());
    }
    op->graphs.push_back({q, psi, g0, G, cfg.mode, cfg.block, cfg.per_cu, order, exec});
  }
  LB_HIP(hipEventRecord(op->ev0, op->stream));
  LB_HIP(hipGraphLaunch(exec, op->stream));
  LB_HIP(hipEventRecord(op->ev1, op->stream));
  return LB_OK;
}

// One sweep issued in a fixed launch configuration.  order: kOrderNatural, or
// (one group of a single problem) kOrderFlux, the flux stored in sweep order
// from a natural source (out of place), or kOrderBoth, source and flux in
// sweep order (in place allowed).
static int launch_sweep_as(lb_op* op, const double* q, double* psi, int g0, int G, const SweepConfig& cfg,
                           int order = kOrderNatural, float* psi32 = nullptr) {
  const int mode = cfg.mode, block = cfg.block;
  const bool pos_out = order != kOrderNatural;
  if (order == kOrderFlux32 && (mode != LB_SWEEP_DIRECTION_LDS || !psi32)) {
    set_error("sweep: the FP32 flux output needs the LDS sweep and an FP32 vector");
    return LB_E_ARG;
  }
  if (pos_out && (G != 1 || (op->batch.P > 0) != (mode == LB_SWEEP_PERSISTENT) || (order == kOrderFlux && q == psi))) {
    set_error("sweep: a sweep-order flux is one group, swept by the persistent kernel on a problem set and by the "
              "other kernels otherwise, out of place unless the source is in sweep order too");
    return LB_E_ARG;
  }
  if ((mode == LB_SWEEP_DIRECTION || mode == LB_SWEEP_DIRECTION_LDS || mode == LB_SWEEP_PERSISTENT) &&
      !((block == 256 || block == 512 || block == 1024 || (block == 768 && mode != LB_SWEEP_PERSISTENT)) &&
        block / G >= 1)) {
    set_error("sweep: the workgroup size must be 256, 512, 768 (direction-resident only) or 1024 and hold the "
              "group batch");
    return LB_E_ARG;
  }
  if (mode == LB_SWEEP_PERSISTENT) {
    if (op->batch.P == 0) {
      set_error("sweep: the persistent sweep needs the problem set (lb_op_set_problems)");
      return LB_E_ARG;
    }
    if (cfg.per_cu < 1 || cfg.per_cu > 8 || block * cfg.per_cu > 2048) {
      set_error("sweep: persistent workgroups per CU must be 1..8 with at most 2048 threads per CU");
      return LB_E_ARG;
    }
    BatchTable& B = op->batch;
    if (pos_out) {
      LB_TRY(ensure_inverse(op));
      LB_TRY(ensure_records(op, B.d_rec_pos, B.d_tasks, B.d_inv.as<const int32_t>()));
    } else {
      LB_TRY(ensure_records(op, B.d_rec, B.d_tasks));
    }
    SweepArgs A{(pos_out ? B.d_rec_pos : B.d_rec).as<const TaskRec>(), op->d_sigt.as<const double>(),
                op->G, g0, G, op->NE, q, psi};
    A.vec1 = one_group_vectors(G, psi);
    A.pos_out = pos_out ? 1 : 0;
    A.pos_in = order == kOrderBoth ? 1 : 0;
    const bool nt = G >= kStreamGroups;
    const int4* items = B.d_items.as<const int4>();
    const int nitems = (int)(B.items.size() / 4);
    const int32_t* lvl = B.d_lvl.as<const int32_t>();
    const int32_t* act = B.d_active.as<const int32_t>();
    unsigned int* ctr = B.d_counter.as<unsigned int>();
    const dim3 grid((unsigned)(B.ncu * cfg.per_cu));
    return replay_graph(op, q, psi, g0, G, cfg, order, [&]() {
      hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, op->stream, ctr);
      launch_persistent(block, nt, grid, op->stream, items, nitems, lvl, act, ctr, A);
    });
  }

  if (pos_out) {
    LB_TRY(ensure_inverse(op));
    LB_TRY(ensure_records(op, op->d_rec_pos, op->d_tasks, order_inverse(op)));
  } else {
    LB_TRY(ensure_records(op, op->d_rec, op->d_tasks));
  }
  SweepArgs A{(pos_out ? op->d_rec_pos : op->d_rec).as<const TaskRec>(), op->d_sigt.as<const double>(),
              op->G, g0, G, op->NE, q, psi};
  A.vec1 = one_group_vectors(G, psi);
  A.pos_out = pos_out ? 1 : 0;
  A.pos_in = order == kOrderBoth ? 1 : 0;
  A.psi32 = psi32;
  const bool nt = G >= kStreamGroups;
  if (mode == LB_SWEEP_DIRECTION || mode == LB_SWEEP_DIRECTION_LDS) {
    if (!op->d_lvl_off.p) {
      LB_TRY(op->d_lvl_off.reserve(op->h_lvl_off.size() * sizeof(int32_t)));
      LB_HIP(hipMemcpy(op->d_lvl_off.p, op->h_lvl_off.data(), op->h_lvl_off.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice));
      LB_TRY(op->d_nlev.reserve(op->h_nlev.size() * sizeof(int32_t)));
      LB_HIP(hipMemcpy(op->d_nlev.p, op->h_nlev.data(), op->h_nlev.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    const int32_t* lo = op->d_lvl_off.as<const int32_t>();
    const int32_t* nl = op->d_nlev.as<const int32_t>();
    if (mode == LB_SWEEP_DIRECTION_LDS) {
      int rc = LB_OK;
      TracePlan* plan = trace_plan(op, G, &rc);
      if (!plan) return rc;
      if (pos_out) LB_TRY(ensure_stage_positions(op, plan));
      const size_t bytes = (size_t)(plan->ring + 2 * plan->stage) * 4 * G * sizeof(double);
      A.uslot = plan->d_uslot.as<const short4>();
      A.wslot = plan->d_wslot.as<const int16_t>();
      const int32_t* so = plan->d_stage_off.as<const int32_t>();
      A.gprod = plan->d_gprod.as<const uint8_t>();
      A.stage = (pos_out ? plan->d_stage_pos : plan->d_stage).as<const int2>();
      LB_HIP(hipEventRecord(op->ev0, op->stream));
      const unsigned nd = (unsigned)op->NI;
      if (nt)
        LB_TRY((launch_direction_lds_nt<true>(block, nd, op->stream, lo, nl, op->maxlev, so, A, bytes)));
      else if (order == kOrderFlux32)
        LB_TRY((launch_direction_lds_nt<false, true>(block, nd, op->stream, lo, nl, op->maxlev, so, A, bytes)));
      else
        LB_TRY((launch_direction_lds_nt<false>(block, nd, op->stream, lo, nl, op->maxlev, so, A, bytes)));
      LB_HIP(hipGetLastError());
      LB_HIP(hipEventRecord(op->ev1, op->stream));
      return LB_OK;
    }
    LB_HIP(hipEventRecord(op->ev0, op->stream));
    launch_direction(block, nt, (unsigned)op->NI, op->stream, lo, nl, op->maxlev, A);
    LB_HIP(hipGetLastError());
    LB_HIP(hipEventRecord(op->ev1, op->stream));
    return LB_OK;
  }

  int rc = LB_OK;
  LevelTable* t = level_table(op, G, &rc);
  if (!t) return rc;
  const int4* work = t->d_entries.as<const int4>();
  auto enqueue = [&]() {
    for (int l = 0; l < op->maxlev; ++l) {
      const int nb = t->level_off[l + 1] - t->level_off[l];
      if (nb == 0) continue;
      if (nt)
        hipLaunchKernelGGL(sweep_level_kernel<true>, dim3(nb), dim3(kBlock), 0, op->stream, work + t->level_off[l], A);
      else
        hipLaunchKernelGGL(sweep_level_kernel<false>, dim3(nb), dim3(kBlock), 0, op->stream, work + t->level_off[l], A);
    }
  };
  if (mode == LB_SWEEP_GRAPH) return replay_graph(op, q, psi, g0, G, cfg, order, enqueue);
  LB_HIP(hipEventRecord(op->ev0, op->stream));
  enqueue();
  LB_HIP(hipGetLastError());
  LB_HIP(hipEventRecord(op->ev1, op->stream));
  return LB_OK;
}

int launch_delta_down(lb_op* op, int64_t n, double* x) {
  hipLaunchKernelGGL(delta_down_kernel, dim3(grid_stride(n)), dim3(kBlock), 0, op->stream, n, op->NENK,
                     op->d_mat.as<const int32_t>(), op->d_xsdd.as<const double>(), x);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_sweep_direction(lb_op* op, const double* q, double* psi, int g0, int G, int block) {
  if (G < 1 || G > block) {
    set_error("sweep: the group batch does not fit the workgroup");
    return LB_E_ARG;
  }
  return launch_sweep_as(op, q, psi, g0, G, SweepConfig{LB_SWEEP_DIRECTION, block, 0});
}

// A sweep in the operator's mode; in LB_SWEEP_AUTO the configuration is timed
// on the first sweeps of each group-batch size (separately for sweep-order
// output) and then fixed.
static int launch_sweep_tuned(lb_op* op, const double* q, double* psi, int g0, int G, int order,
                              float* psi32 = nullptr) {
  if (op->sweep_mode != LB_SWEEP_AUTO && order != kOrderFlux32)
    return launch_sweep_as(op, q, psi, g0, G, SweepConfig{op->sweep_mode, op->dir_block, op->per_cu}, order);
  if (op->sweep_mode == LB_SWEEP_DIRECTION_LDS)
    return launch_sweep_as(op, q, psi, g0, G, SweepConfig{op->sweep_mode, op->dir_block, op->per_cu}, order, psi32);
  const bool batched = op->batch.P > 0;
  const SweepConfig* table = batched ? kBatchTuneConfig : kTuneConfig;
  const int nconf = batched ? kBatchTuneConfigs : kTuneConfigs;
  SweepTune& tn = order != kOrderNatural ? (batched ? op->btune_pos[order] : op->tune_pos[order])
                                         : (batched ? op->btune[G] : op->tune[G]);
  if (tn.choice >= 0) return launch_sweep_as(op, q, psi, g0, G, table[tn.choice], order, psi32);
  // Timing phase: each configuration sweeps kTuneTrials times (round robin), each
  // sweep timed alone between the operator's events; configurations whose block
  // cannot hold the group batch are skipped, and so is the LDS sweep with many
  // groups per cell, whose LDS then holds only a few hundred cells (24 groups:
  // 200 ms against 53 ms per sweep, plus a second slot plan).  Results are
  // identical in every mode.
  // (a sweep-order flux of a problem set follows the persistent kernel's order, so
  // only its configurations are timed there)
  auto skip = [&](const SweepConfig& k) {
    return (k.block != 0 && k.block < G) || (k.mode == LB_SWEEP_DIRECTION_LDS && G >= kStreamGroups) ||
           (batched && order != kOrderNatural && k.mode != LB_SWEEP_PERSISTENT) ||
           (order == kOrderFlux32 && k.mode != LB_SWEEP_DIRECTION_LDS);
  };
  int c = tn.trials % nconf;
  while (skip(table[c])) {
    ++tn.trials;
    c = tn.trials % nconf;
  }
  LB_TRY(launch_sweep_as(op, q, psi, g0, G, table[c], order, psi32));
  LB_HIP(hipEventSynchronize(op->ev1));
  float ms = 0.f;
  LB_HIP(hipEventElapsedTime(&ms, op->ev0, op->ev1));
  tn.best_ms[c] = std::min(tn.best_ms[c], ms);
  if (++tn.trials >= kTuneTrials * nconf) {
    int best = 0;
    for (int i = 1; i < nconf; ++i)
      if (tn.best_ms[i] < tn.best_ms[best]) best = i;
    tn.choice = best;
  }
  return LB_OK;
}

int launch_sweep(lb_op* op, const double* q, double* psi, int g0, int G) {
  if (G < 1 || G > kBlock) {
    set_error("sweep: the group batch must be between 1 and 256");
    return LB_E_ARG;
  }
  return launch_sweep_tuned(op, q, psi, g0, G, kOrderNatural);
}

int launch_sweep_pos(lb_op* op, const double* q, double* psi_so, int g0, bool source_in_order) {
  return launch_sweep_tuned(op, q, psi_so, g0, 1, source_in_order ? kOrderBoth : kOrderFlux);
}

int launch_sweep_f32(lb_op* op, const double* q, double* psi_so, float* psi32_so, int g0) {
  if (op->batch.P > 0) {
    set_error("sweep: the FP32 flux output is for single problems");
    return LB_E_ARG;
  }
  if ((reinterpret_cast<uintptr_t>(psi_so) & 15) || (reinterpret_cast<uintptr_t>(psi32_so) & 15)) {
    set_error("sweep: the FP32 flux output needs 16-byte aligned flux vectors");   // each cell: 16-byte stores
    return LB_E_ARG;
  }
  return launch_sweep_tuned(op, q, psi_so, g0, 1, kOrderFlux32, psi32_so);
}

bool sweep_order_fp32(const lb_op* op) {
  return op->flux_order == LB_FLUX_SWEEP_ORDER_FP32 && op->nmom <= 16 && op->moment_fmt == LB_FMT_FP64 &&
         op->batch.P == 0;
}

bool sweep_order_flux(const lb_op* op) {
  return op->flux_order != LB_FLUX_NATURAL && op->nmom <= 16 && op->moment_fmt == LB_FMT_FP64 &&
         (op->batch.P == 0 || op->sweep_mode == LB_SWEEP_AUTO || op->sweep_mode == LB_SWEEP_PERSISTENT);
}

int launch_flux_to_natural(lb_op* op, const double* psi_so, double* psi) {
  LB_TRY(ensure_inverse(op));
  const int64_t n = op->NI * op->NE * 2;
  hipLaunchKernelGGL(flux_to_natural_kernel, dim3(grid_for(n)), dim3(kBlock), 0, op->stream, (int)op->NI, op->NE,
                     order_inverse(op), psi_so, psi);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

#define LB_NM_DISPATCH(NMV, CALL) \
  switch (NMV) {                  \
    case 1: CALL(1); break;       \
    case 4: CALL(4); break;       \
    case 9: CALL(9); break;       \
    case 16: CALL(16); break;     \
    case 25: CALL(25); break;     \
    case 36: CALL(36); break;     \
    case 49: CALL(49); break;     \
    case 64: CALL(64); break;     \
    default:                      \
      set_error("moments: moment count is not (L+1)^2"); \
      return LB_E_ARG;            \
  }

#define LB_NM_DISPATCH16(NMV, CALL) \
  switch (NMV) {                    \
    case 1: CALL(1); break;         \
    case 4: CALL(4); break;         \
    case 9: CALL(9); break;         \
    case 16: CALL(16); break;       \
    default:                        \
      set_error("moments: moment count is not (L+1)^2"); \
      return LB_E_ARG;              \
  }

// Moment counts above this go through the chunked kernels when rocBLAS is
// not used.
constexpr int kMaxRegisterMoments = 64;

// When the D and M products are rocBLAS FP64 GEMMs: from L = 4 on, where they
// are compute-bound (2 nmom NI flops per column against 8 (NI + nmom) bytes) and
// the matrix cores win.  Up to L = 3 they are HBM-bound and the two-column
// kernels above stream faster than the GEMM.
static bool use_blas(const lb_op* op) { return op->nmom >= 25; }

// The two-column kernels need an even column count (always: 4 NE G) and
// 16-byte aligned vectors.
static bool pair_ok(int64_t ncols, const void* a, const void* b) {
  return ncols % 2 == 0 && (reinterpret_cast<uintptr_t>(a) & 15) == 0 && (reinterpret_cast<uintptr_t>(b) & 15) == 0;
}

static rocblas_handle blas_handle(lb_op* op, int* rc) {
  if (!op->blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) {
      set_error("moments: rocblas_create_handle failed");
      *rc = LB_E_HIP;
      return nullptr;
    }
    op->blas = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(op->blas);
  if (rocblas_set_stream(h, op->stream) != rocblas_status_success) {
    set_error("moments: rocblas_set_stream failed");
    *rc = LB_E_HIP;
    return nullptr;
  }
  *rc = LB_OK;
  return h;
}

void release_blas(lb_op* op) {
  if (op->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(op->blas));
  op->blas = nullptr;
}

// phi (ncols x nmom, column-major, = phi[lm][col]) = psi (ncols x NI, = psi[i][col])
// times D^T (NI x nmom, column-major with leading dimension NI, = D[lm][i]).
static int blas_moments(lb_op* op, const double* psi, double* phi, int64_t ncols) {
  int rc = LB_OK;
  rocblas_handle h = blas_handle(op, &rc);
  if (!h) return rc;
  const double one = 1.0, zero = 0.0;
  if (rocblas_dgemm_64(h, rocblas_operation_none, rocblas_operation_none, ncols, op->nmom, op->NI, &one, psi, ncols,
                       op->d_D.as<const double>(), op->NI, &zero, phi, ncols) != rocblas_status_success) {
    set_error("moments: rocblas_dgemm_64 (D psi) failed");
    return LB_E_HIP;
  }
  return LB_OK;
}

// psi (ncols x NI) (+)= phis (ncols x nmom) times M^T (nmom x NI, column-major with
// leading dimension nmom, = M[i][lm]).
static int blas_moments_to_discrete(lb_op* op, const double* phis, double* psi, int64_t ncols, bool accumulate) {
  int rc = LB_OK;
  rocblas_handle h = blas_handle(op, &rc);
  if (!h) return rc;
  const double one = 1.0, beta = accumulate ? 1.0 : 0.0;
  if (rocblas_dgemm_64(h, rocblas_operation_none, rocblas_operation_none, ncols, op->NI, op->nmom, &one, phis, ncols,
                       op->d_M.as<const double>(), op->nmom, &beta, psi, ncols) != rocblas_status_success) {
    set_error("moments: rocblas_dgemm_64 (M phi) failed");
    return LB_E_HIP;
  }
  return LB_OK;
}

int launch_moments(lb_op* op, const double* psi, double* phi, int G, int64_t ncols) {
  (void)G;
  if (use_blas(op)) return blas_moments(op, psi, phi, ncols);
  if (op->nmom > kMaxRegisterMoments) {
    const unsigned chunks = (unsigned)((op->nmom + kMomChunk - 1) / kMomChunk);
    hipLaunchKernelGGL(moments_chunk_kernel, dim3(grid_for(ncols), chunks), dim3(kBlock), 0, op->stream,
                       op->d_D.as<const double>(), (int)op->NI, (int)op->nmom, ncols, psi, phi);
    LB_HIP(hipGetLastError());
    return LB_OK;
  }
  if (op->nmom <= 16 && pair_ok(ncols, psi, phi)) {
#define CALL(NMV)                                                                                                   \
  hipLaunchKernelGGL(moments_pair_kernel<NMV>, dim3(grid_for(ncols / 2)), dim3(kBlock), 0, op->stream,              \
                     op->d_D.as<double>(), (int)op->NI, ncols, psi, phi)
    LB_NM_DISPATCH16(op->nmom, CALL)
#undef CALL
    LB_HIP(hipGetLastError());
    return LB_OK;
  }
#define CALL(NMV)                                                                                                  \
  hipLaunchKernelGGL(moments_kernel<NMV>, dim3(grid_for(ncols)), dim3(kBlock), 0, op->stream, op->d_D.as<double>(), \
                     (int)op->NI, ncols, psi, phi)
  LB_NM_DISPATCH(op->nmom, CALL)
#undef CALL
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_moments_pos(lb_op* op, const double* psi_so, double* phi) {
  if (op->nmom > 16 || !pair_ok(op->NENK, psi_so, phi)) {
    set_error("moments: a sweep-order flux takes at most 16 moments and 16-byte aligned vectors");
    return LB_E_ARG;
  }
  LB_TRY(ensure_inverse(op));
#define CALL(NMV)                                                                                                   \
  hipLaunchKernelGGL(moments_gather_pair_kernel<NMV>, dim3(grid_for(op->NENK / 2)), dim3(kBlock), 0, op->stream,     \
                     op->d_D.as<const double>(), (int)op->NI, op->NE, order_inverse(op), psi_so, phi)
  LB_NM_DISPATCH16(op->nmom, CALL)
#undef CALL
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_moments_pos32(lb_op* op, const float* psi32_so, double* phi) {
  if (op->nmom > 16 || !pair_ok(op->NENK, psi32_so, phi)) {
    set_error("moments: an FP32 sweep-order flux takes at most 16 moments and 16-byte aligned vectors");
    return LB_E_ARG;
  }
  LB_TRY(ensure_inverse(op));
#define CALL(NMV)                                                                                                   \
  hipLaunchKernelGGL(moments_gather32_pair_kernel<NMV>, dim3(grid_for(op->NENK / 2)), dim3(kBlock), 0, op->stream,   \
                     op->d_D.as<const double>(), (int)op->NI, op->NE, order_inverse(op), psi32_so, phi)
  LB_NM_DISPATCH16(op->nmom, CALL)
#undef CALL
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_moments_to_order(lb_op* op, const double* phi, double* psi_so, int g0, bool scaled) {
  if (op->nmom > 16 || !pair_ok(op->NENK, phi, psi_so)) {
    set_error("moments: a sweep-order flux takes at most 16 moments and 16-byte aligned vectors");
    return LB_E_ARG;
  }
  LB_TRY(ensure_inverse(op));
  const double* sigs = scaled ? op->d_sigs.as<const double>() : nullptr;
#define CALL(NMV)                                                                                                   \
  hipLaunchKernelGGL(moments_to_order_pair_kernel<NMV>, dim3(grid_for(op->NENK / 2)), dim3(kBlock), 0, op->stream,   \
                     op->d_M.as<const double>(), op->d_deg.as<const int32_t>(), (int)op->NI, op->NE, op->G, g0,      \
                     op->L, op->d_mat.as<const int32_t>(), sigs, order_inverse(op), phi, psi_so)
  LB_NM_DISPATCH16(op->nmom, CALL)
#undef CALL
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_moments_to_discrete(lb_op* op, const double* phi, double* psi, int g0, int G, int accumulate, bool scaled) {
  const int64_t ncols = op->NENK * G;
  const double* sigs = scaled ? op->d_sigs.as<const double>() : nullptr;
  if (use_blas(op)) {
    const double* in = phi;
    if (sigs) {
      LB_TRY(op->w_phis.reserve((size_t)op->nmom * ncols * sizeof(double)));
      hipLaunchKernelGGL(scale_moments_kernel, dim3(grid_stride((int64_t)op->nmom * ncols)), dim3(kBlock), 0,
                         op->stream, phi, (int)op->nmom, ncols, G, op->G, g0, op->L, op->d_deg.as<const int32_t>(),
                         op->d_mat.as<const int32_t>(), sigs, op->w_phis.as<double>());
      LB_HIP(hipGetLastError());
      in = op->w_phis.as<const double>();
    }
    return blas_moments_to_discrete(op, in, psi, ncols, accumulate != 0);
  }
  if (op->nmom > kMaxRegisterMoments) {
    for (int m0 = 0; m0 < op->nmom; m0 += kMomChunk)
      hipLaunchKernelGGL(moments_to_discrete_chunk_kernel, dim3(grid_for(ncols)), dim3(kBlock), 0, op->stream,
                         op->d_M.as<const double>(), op->d_deg.as<const int32_t>(), (int)op->NI, (int)op->nmom, m0,
                         ncols, G, op->G, g0, op->L, op->d_mat.as<const int32_t>(), sigs, phi, psi,
                         (m0 > 0 || accumulate) ? 1 : 0);
    LB_HIP(hipGetLastError());
    return LB_OK;
  }
  if (op->nmom <= 16 && pair_ok(ncols, phi, psi)) {
#define CALL(NMV)                                                                                                 \
  hipLaunchKernelGGL(moments_to_discrete_pair_kernel<NMV>, dim3(grid_for(ncols / 2)), dim3(kBlock), 0, op->stream, \
                     op->d_M.as<double>(), op->d_deg.as<const int32_t>(), (int)op->NI, ncols, G, op->G, g0, op->L,  \
                     op->d_mat.as<const int32_t>(), sigs, phi, psi, accumulate)
    LB_NM_DISPATCH16(op->nmom, CALL)
#undef CALL
    LB_HIP(hipGetLastError());
    return LB_OK;
  }
#define CALL(NMV)                                                                                                 \
  hipLaunchKernelGGL(moments_to_discrete_kernel<NMV>, dim3(grid_for(ncols)), dim3(kBlock), 0, op->stream,       \
                     op->d_M.as<double>(), op->d_deg.as<const int32_t>(), (int)op->NI, ncols, G, op->G, g0, op->L,             \
                     op->d_mat.as<const int32_t>(), sigs, phi, psi, accumulate)
  LB_NM_DISPATCH(op->nmom, CALL)
#undef CALL
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_scatter_dense(lb_op* op, const double* x, double* y, int g0, int G) {
  (void)g0;
  const int64_t ncols = op->NENK * G;
  hipLaunchKernelGGL(scatter_dense_kernel, dim3(grid_for(ncols)), dim3(kBlock), 0, op->stream, op->d_yys.as<double>(),
                     (int)op->NI, ncols, G, op->d_mat.as<const int32_t>(), x, y);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_downscatter(lb_op* op, const double* phi_all, const double* S, int g, int G, const double* qmom,
                       double* acc) {
  const int64_t tot = (int64_t)op->nmom * op->NENK;
  hipLaunchKernelGGL(downscatter_kernel, dim3(grid_for(tot)), dim3(kBlock), 0, op->stream, phi_all, S, g, G, op->Nm,
                     op->nmom, op->L, op->d_mat.as<const int32_t>(), op->d_deg.as<const int32_t>(), op->NENK, qmom,
                     acc);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_axpby(hipStream_t s, int64_t n, double a, const double* x, double b, const double* y, double* z) {
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_stride(n)), dim3(kBlock), 0, s, n, a, x, b, y, z);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_scale(hipStream_t s, int64_t n, double a, double* x) {
  hipLaunchKernelGGL(scale_kernel, dim3(grid_stride(n)), dim3(kBlock), 0, s, n, a, x);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int launch_copy(hipStream_t s, int64_t n, const double* x, double* y) {
  if (x == y || n == 0) return LB_OK;
  LB_HIP(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  return LB_OK;
}

int launch_dot(hipStream_t s, int64_t n, const double* x, const double* y, double* part, double* res) {
  hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(kBlock), 0, s, n, x, y, part);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kBlock), 0, s, kDotBlocks, part, res);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int dot(lb_op* op, int64_t n, const double* x, const double* y, double* out) {
  LB_TRY(op->w_red.reserve((kDotBlocks + 8) * sizeof(double)));
  if (!op->w_host.p) {
    LB_HIP(hipHostMalloc(&op->w_host.p, kMaxGroupBatch * sizeof(double), hipHostMallocDefault));
    op->w_host.bytes = kMaxGroupBatch * sizeof(double);
  }
  double* part = op->w_red.as<double>();
  double* res = part + kDotBlocks;
  LB_TRY(launch_dot(op->stream, n, x, y, part, res));
  LB_HIP(hipMemcpyAsync(op->w_host.p, res, sizeof(double), hipMemcpyDeviceToHost, op->stream));
  LB_HIP(hipStreamSynchronize(op->stream));
  *out = *static_cast<double*>(op->w_host.p);
  return LB_OK;
}

int multi_dot(lb_op* op, int64_t n, const double* V, int64_t ld, int c, const double* w, bool self, double* out) {
  const int slots = c + (self ? 1 : 0);
  if (c < 0 || slots < 1) {
    set_error("multi_dot: no vectors");
    return LB_E_ARG;
  }
  LB_TRY(op->w_red.reserve((size_t)slots * (kOrthoBlocks + 1) * sizeof(double)));
  if (!op->w_host.p || op->w_host.bytes < slots * sizeof(double)) {
    if (op->w_host.p) (void)hipHostFree(op->w_host.p);
    op->w_host.p = nullptr;
    const size_t want = std::max<size_t>(slots, kMaxGroupBatch) * sizeof(double);
    LB_HIP(hipHostMalloc(&op->w_host.p, want, hipHostMallocDefault));
    op->w_host.bytes = want;
  }
  double* part = op->w_red.as<double>();
  double* res = part + (size_t)slots * kOrthoBlocks;
  for (int i0 = 0; i0 < c || (i0 == 0 && self); i0 += kOrthoChunk) {
    const int cc = std::min(kOrthoChunk, c - i0);
    hipLaunchKernelGGL(multi_dot_partial_kernel, dim3(kOrthoBlocks), dim3(kBlock), 0, op->stream, n, cc,
                       V + (int64_t)i0 * ld, ld, w, i0, (self && i0 == 0) ? c : -1, part, nullptr, 0);
  }
  hipLaunchKernelGGL(multi_dot_final_kernel, dim3(slots), dim3(kBlock), 0, op->stream, part, res, nullptr, 0);
  LB_HIP(hipGetLastError());
  LB_HIP(hipMemcpyAsync(op->w_host.p, res, slots * sizeof(double), hipMemcpyDeviceToHost, op->stream));
  LB_HIP(hipStreamSynchronize(op->stream));
  std::copy(static_cast<double*>(op->w_host.p), static_cast<double*>(op->w_host.p) + slots, out);
  return LB_OK;
}

int multi_axpy(lb_op* op, int64_t n, const double* V, int64_t ld, int c, const double* h, double scale, double* w) {
  if (c == 0) return scale == 1.0 ? LB_OK : launch_scale(op->stream, n, scale, w);
  for (int i0 = 0; i0 < c; i0 += kOrthoChunk) {
    const int cc = std::min(kOrthoChunk, c - i0);
    OrthoCoef k{};
    for (int i = 0; i < cc; ++i) k.h[i] = h[i0 + i];
    hipLaunchKernelGGL(multi_axpy_kernel, dim3(grid_stride(n)), dim3(kBlock), 0, op->stream, n, cc,
                       V + (int64_t)i0 * ld, ld, k, i0 + cc == c ? scale : 1.0, w);
  }
  LB_HIP(hipGetLastError());
  return LB_OK;
}

namespace {

// res[i] = <w, V_i> (i < c), res[c] = <w, w>, left in device memory; the
// kernels run only when the gate is open.
int multi_dot_enqueue(lb_op* op, int64_t n, const double* V, int64_t ld, int c, const double* w, double* res,
                      const int* gate, int mask) {
  const int slots = c + 1;
  LB_TRY(op->w_red.reserve((size_t)slots * kOrthoBlocks * sizeof(double)));
  double* part = op->w_red.as<double>();
  for (int i0 = 0; i0 < c || i0 == 0; i0 += kOrthoChunk) {
    const int cc = std::min(kOrthoChunk, c - i0);
    hipLaunchKernelGGL(multi_dot_partial_kernel, dim3(kOrthoBlocks), dim3(kBlock), 0, op->stream, n, cc,
                       V + (int64_t)i0 * ld, ld, w, i0, i0 == 0 ? c : -1, part, gate, mask);
  }
  hipLaunchKernelGGL(multi_dot_final_kernel, dim3(slots), dim3(kBlock), 0, op->stream, part, res, gate, mask);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

// w <- (w - sum_{i<c} coef_i V_i) * (*scale_p, or 1), coefficients in device memory.
int multi_axpy_enqueue(lb_op* op, int64_t n, const double* V, int64_t ld, int c, const double* coef,
                       const double* scale_p, const int* gate, int mask, double* w) {
  for (int i0 = 0; i0 < c; i0 += kOrthoChunk) {
    const int cc = std::min(kOrthoChunk, c - i0);
    hipLaunchKernelGGL(multi_axpy_dev_kernel, dim3(grid_stride(n)), dim3(kBlock), 0, op->stream, n, cc,
                       V + (int64_t)i0 * ld, ld, coef + i0, i0 + cc == c ? scale_p : nullptr, gate, mask, w);
  }
  LB_HIP(hipGetLastError());
  return LB_OK;
}

}  // namespace

int arnoldi_reserve(lb_op* op, int m, ArnoldiDev* a) {
  if (m < 1) {
    set_error("arnoldi: restart length must be positive");
    return LB_E_ARG;
  }
  const size_t ldh = (size_t)m + 1;
  const size_t doubles = ldh * m + m + m + ldh + ldh + (ldh + 1) + ldh + m;
  LB_TRY(op->w_arn.reserve(sizeof(ArnoldiCtl) + doubles * sizeof(double)));
  a->m = m;
  a->ctl = op->w_arn.as<ArnoldiCtl>();
  double* p = reinterpret_cast<double*>(a->ctl + 1);
  a->H = p;
  p += ldh * m;
  a->cs = p;
  p += m;
  a->sn = p;
  p += m;
  a->g = p;
  p += ldh;
  a->hk = p;
  p += ldh;
  a->d = p;
  p += ldh + 1;
  a->coef = p;
  p += ldh;
  a->y = p;
  if (!op->w_host.p || op->w_host.bytes < sizeof(ArnoldiCtl)) {
    if (op->w_host.p) (void)hipHostFree(op->w_host.p);
    op->w_host.p = nullptr;
    const size_t want = std::max<size_t>(sizeof(ArnoldiCtl), kMaxGroupBatch * sizeof(double));
    LB_HIP(hipHostMalloc(&op->w_host.p, want, hipHostMallocDefault));
    op->w_host.bytes = want;
  }
  return LB_OK;
}

int arnoldi_start(lb_op* op, const ArnoldiDev& a, double beta) {
  hipLaunchKernelGGL(arnoldi_start_kernel, dim3(1), dim3(64), 0, op->stream, a.g, a.m, beta);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

int arnoldi_step(lb_op* op, int64_t n, const double* V, int k, const ArnoldiDev& a, bool shifted, ArnoldiCtl* out) {
  if (k < 0 || k >= a.m) {
    set_error("arnoldi: step outside the cycle");
    return LB_E_ARG;
  }
  hipStream_t s = op->stream;
  const int c = k + 1;
  double* w = const_cast<double*>(V) + (int64_t)c * n;
  const int* st = &a.ctl->state;
  LB_TRY(multi_dot_enqueue(op, n, V, n, c, w, a.d, nullptr, 0));
  hipLaunchKernelGGL(ortho_decide_kernel, dim3(1), dim3(64), 0, s, a.d, c, 0, a.hk, a.coef, a.ctl);
  if (c <= kOrthoChunk) {   // the update and the second pass's dots in one read of V and w
    hipLaunchKernelGGL(multi_axpy_dot_kernel, dim3(kOrthoBlocks), dim3(kBlock), 0, s, n, c, V, n, a.coef,
                       &a.ctl->scale, st, (1 << kArnDone1) | (1 << kArnPass2), 1 << kArnPass2, w,
                       op->w_red.as<double>());
    hipLaunchKernelGGL(multi_dot_final_kernel, dim3(c + 1), dim3(kBlock), 0, s, op->w_red.as<const double>(), a.d,
                       st, 1 << kArnPass2);
  } else {
    LB_TRY(multi_axpy_enqueue(op, n, V, n, c, a.coef, &a.ctl->scale, st, (1 << kArnDone1) | (1 << kArnPass2), w));
    LB_TRY(multi_dot_enqueue(op, n, V, n, c, w, a.d, st, 1 << kArnPass2));
  }
  hipLaunchKernelGGL(ortho_decide_kernel, dim3(1), dim3(64), 0, s, a.d, c, 1, a.hk, a.coef, a.ctl);
  LB_TRY(multi_axpy_enqueue(op, n, V, n, c, a.coef, &a.ctl->scale, st, 1 << kArnDone2, w));
  hipLaunchKernelGGL(arnoldi_givens_kernel, dim3(1), dim3(64), 0, s, a.H, a.cs, a.sn, a.g, a.hk, k, a.m + 1,
                     shifted ? 1 : 0, a.ctl);
  LB_HIP(hipGetLastError());
  LB_HIP(hipMemcpyAsync(op->w_host.p, a.ctl, sizeof(ArnoldiCtl), hipMemcpyDeviceToHost, s));
  LB_HIP(hipStreamSynchronize(s));
  std::memcpy(out, op->w_host.p, sizeof(ArnoldiCtl));
  return LB_OK;
}

int arnoldi_update(lb_op* op, int64_t n, const double* V, int k, const ArnoldiDev& a, double* sol) {
  if (k == 0) return LB_OK;
  hipLaunchKernelGGL(arnoldi_solve_kernel, dim3(1), dim3(64), 0, op->stream, a.H, a.g, k, a.m + 1, a.y);
  LB_HIP(hipGetLastError());
  return multi_axpy_enqueue(op, n, V, n, k, a.y, nullptr, nullptr, 0, sol);
}

int launch_fluence(lb_op* op, const double* v, double* phi) {
  hipLaunchKernelGGL(fluence_kernel, dim3(grid_for(op->NENK)), dim3(kBlock), 0, op->stream, op->d_w.as<double>(),
                     (int)op->NI, op->NENK, v, phi);
  LB_HIP(hipGetLastError());
  return LB_OK;
}

}  // namespace lb
