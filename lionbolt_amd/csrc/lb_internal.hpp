// Internal declarations shared by the liblbhip sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/lionbolt.h"

namespace lb {

void set_error(const std::string& msg);

#define LB_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t _e = (call);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      lb::set_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " + __FILE__ + \
                    ":" + std::to_string(__LINE__) + " in " #call);                          \
      return LB_E_HIP;                                                                       \
    }                                                                                        \
  } while (0)

// C entry points run their bodies through guarded(): a C++ exception (a failed
// host allocation of a std::vector, say) becomes a status code instead of
// crossing the C ABI.
template <class F>
int guarded(F&& f) noexcept {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    try {
      set_error("host memory allocation failed");
    } catch (...) {
    }
    return LB_E_NOMEM;
  } catch (const std::exception& e) {
    try {
      set_error(std::string("unexpected C++ exception: ") + e.what());
    } catch (...) {
    }
    return LB_E_ARG;
  } catch (...) {
    return LB_E_ARG;
  }
}

#define LB_TRY(call)              \
  do {                            \
    int _rc = (call);             \
    if (_rc != LB_OK) return _rc; \
  } while (0)

// A grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int reserve(size_t n);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Work list of one sweep configuration: for every global level, the workgroups
// (direction, first sweep position, cell count) that sweep that level.
struct LevelTable {
  int cpw = 0;                        // cells per workgroup
  std::vector<int32_t> level_off;     // [maxlev + 1] offsets into entries
  std::vector<int32_t> entries;       // 4 ints per workgroup
  DevBuf d_entries;
};

// Local faces of a tetrahedron (FiniteElementAnalysis.f90:129-132), 0-based.
constexpr int kFaceNodes[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

struct SweepGraph {
  const double* q;
  double* psi;
  int g0, G, mode, block, per_cu, order;
  hipGraphExec_t exec;
};

// Launch configurations the sweep autotuner times, and its state per group batch.
// A single problem chooses among the graph-replayed level launches and the
// direction-resident kernel; a batch of independent problems among the level
// launches and the persistent work-queue kernel (block, workgroups per CU).
struct SweepConfig {
  int mode, block, per_cu;
};
// Layout of a one-group sweep's vectors: both natural; the flux in sweep order
// from a natural source; source and flux in sweep order.
constexpr int kOrderNatural = 0, kOrderFlux = 1, kOrderBoth = 2;
// The flux in sweep order as FP32 (one group, LDS sweep), from a natural source;
// the FP64 flux only where a later cell reads it from global memory.
constexpr int kOrderFlux32 = 3;
// (The LDS sweep at 1024 threads spills registers and is not timed; at 768 it
// runs three waves per SIMD on one round of look-ahead.)
constexpr int kTuneConfigs = 7;
constexpr SweepConfig kTuneConfig[kTuneConfigs] = {
    {LB_SWEEP_GRAPH, 0, 0},           {LB_SWEEP_DIRECTION, 256, 0},     {LB_SWEEP_DIRECTION, 512, 0},
    {LB_SWEEP_DIRECTION, 1024, 0},    {LB_SWEEP_DIRECTION_LDS, 256, 0}, {LB_SWEEP_DIRECTION_LDS, 512, 0},
    {LB_SWEEP_DIRECTION_LDS, 768, 0}};
constexpr int kBatchTuneConfigs = 8;
constexpr SweepConfig kBatchTuneConfig[kBatchTuneConfigs] = {
    {LB_SWEEP_GRAPH, 0, 0},         {LB_SWEEP_PERSISTENT, 256, 4},  {LB_SWEEP_PERSISTENT, 256, 8},
    {LB_SWEEP_PERSISTENT, 512, 2},  {LB_SWEEP_PERSISTENT, 512, 4},  {LB_SWEEP_PERSISTENT, 1024, 1},
    {LB_SWEEP_PERSISTENT, 1024, 2}, {LB_SWEEP_DIRECTION, 1024, 0}};
constexpr int kTuneTrials = 2;
struct SweepTune {
  int trials = 0;                        // timed sweeps so far (config = trials % number of configs)
  float best_ms[8] = {1e30f, 1e30f, 1e30f, 1e30f, 1e30f, 1e30f, 1e30f, 1e30f};
  int choice = -1;                       // index into the config table once settled
};

// LDS slot plan of the direction-resident LDS sweep for one group-batch size
// (built by lb_trace_plan.cpp on the device sweep order).
struct TracePlan {
  int ring = 0, stage = 0;   // ring slots; staging slots per level parity
  int64_t stats[4] = {0, 0, 0, 0};   // upstream traces read from the ring, staged, one level back, overflow
  DevBuf d_uslot;            // int32 [NI][NE][4]: LDS slot of each upstream trace, -1 = global memory
  DevBuf d_wslot;            // int32 [NI][NE]: slot the cell writes, -1 = none
  DevBuf d_stage_off;        // int32 [NI][maxlev + 2]: first staging entry of each consumer level
  DevBuf d_stage;            // int2 entries: (LDS slot, element whose flux is copied)
  DevBuf d_stage_pos;        // the same entries with the element's sweep position (sweep-order flux)
  DevBuf d_gprod;            // uint8 [NI][NE] (one group): 1 where a consumer reads the cell's flux from
                             // global memory (staged or directly), 0 where only ring slots do
};

// Most energy groups solved together by the group-batched multigroup solve.
constexpr int kMaxGroupBatch = 16;

// Contiguous (problem, row) ranges of a [rows][NENK] vector, split into tiles, for
// the per-problem reductions and updates of batched solves.
struct SegTable {
  std::vector<int64_t> tiles;            // 4 per tile: problem, start, length, 0
  std::vector<int32_t> tile_off;         // [P + 1] tiles of problem p
  DevBuf d_tiles, d_tile_off;
};

// A mesh made of independent problems in consecutive element ranges.
struct BatchTable {
  int P = 0;
  std::vector<int64_t> eoff;             // [P + 1] element offsets
  std::vector<int32_t> pmat;             // material of each problem's first element
  std::vector<int32_t> items;            // 4 per (problem, direction) item: p, i, level base, level count
  std::vector<int32_t> lvl;              // absolute level offsets (sweep positions in batch order)
  std::vector<uint32_t> tasks;           // [NI][NE] task words, problem-major inside each direction
  DevBuf d_items, d_lvl, d_tasks, d_rec, d_active, d_counter, d_coef, d_part, d_out;
  DevBuf d_inv, d_rec_pos;               // sweep-order flux of the persistent sweep: positions, records
  double* h_io = nullptr;                // pinned host staging for per-problem scalars
  std::map<int, SegTable> seg;           // by rows
  int ncu = 0;
};

void clear_graphs(lb_op* op);

}  // namespace lb

struct lb_op {
  int64_t NE = 0, NK = 0, NI = 0, NENK = 0;
  int L = 0, nmom = 1, Nm = 1;
  int maxlev = 0;
  bool host_only = false;           // precompute only (no device): used by lb_schedule_host
  // host-side copies kept for queries and table builds
  std::vector<double> h_nrm;        // [NE][4][3]
  std::vector<double> h_k, h_w;     // [NI][3], [NI]
  std::vector<int32_t> h_mat;       // [NE] 0-based
  std::vector<uint32_t> h_tasks;    // [NI][NE] packed cell | up << 24 | out << 28
  std::vector<int32_t> h_nlev;      // [NI]
  std::vector<int32_t> h_lvl_off;   // [NI][maxlev + 1]
  std::vector<double> h_D, h_M;     // [nmom][NI], [NI][nmom]
  std::vector<double> h_yy;         // [L+1][NI][NI] = YY(ip, i, l) stored [l][ip][i]
  std::vector<int32_t> h_deg;       // [nmom] Legendre order of each moment
  // device geometry and schedule
  lb::DevBuf d_geo;     // double [NE][4][4]: n_x, n_y, n_z, c_f = 10 A_f / detJ
  lb::DevBuf d_nbr;     // int32 [NE][4], -1 on the boundary
  lb::DevBuf d_perm;    // uint32 [NE]: neighbour local node of each face node, 2 bits each
  lb::DevBuf d_mat;     // int32 [NE]
  lb::DevBuf d_tasks;   // uint32 [NI][NE]
  lb::DevBuf d_k;       // double [NI][3]
  lb::DevBuf d_w;       // double [NI]
  lb::DevBuf d_D, d_M;  // moment matrices
  lb::DevBuf d_deg;     // int32 [nmom] Legendre order of each moment
  lb::DevBuf d_yys;     // double [Nm][NI][NI] dense kernel (YYS(ip,i,m) stored [m][ip][i])
  std::map<int, lb::LevelTable> tables;
  // physics
  int G = 0;                        // groups in the physics table
  bool phys = false;
  bool delta_down = false;          // K = XSdd(mat) I (ScatteringOp%BuildDeltaDown)
  lb::DevBuf d_xsdd;                // double [Nm]
  std::vector<double> h_sigt;       // [Nm][G]
  std::vector<double> h_sigs;       // [Nm][G][L+1]
  lb::DevBuf d_sigt, d_sigs;
  int scatter_mode = LB_SCATTER_MOMENT;
  int sweep_mode = LB_SWEEP_AUTO;
  std::map<int, lb::SweepTune> tune;  // autotuner state per group-batch size
  std::map<int, lb::SweepTune> tune_pos;  // the same for one-group sweeps in sweep order, by lb::kOrder*
  int flux_order = LB_FLUX_SWEEP_ORDER;   // layout of the within-group solves' angular flux
  lb::DevBuf d_inv;                 // int32 [NI][NE]: sweep position of each element per direction
  lb::DevBuf d_rec_pos;             // d_rec with the neighbours' sweep positions in place of their elements
  lb::DevBuf d_rec;                 // per (direction, sweep position): task word, permutation and material,
                                    // neighbours, c_f (k . n_f) of three faces (48 B)
  lb::DevBuf d_lvl_off, d_nlev;     // level offsets per direction, for the direction-resident sweep
  std::map<int, lb::TracePlan> plans;  // LDS sweep slot plans per group-batch size (built on first use)
  int lds_max = 0;                  // LDS bytes a workgroup may allocate (queried once)
  void* blas = nullptr;             // rocBLAS handle of the FP64 moment GEMMs (created on first use)
  lb::DevBuf w_phis;                // sigma-scaled flux moments, input of the M GEMM
  int dir_block = 1024;             // workgroup size of the direction-resident and persistent sweeps
  int per_cu = 2;                   // persistent sweep: resident workgroups per compute unit
  int group_batch = 1;              // energy groups per coupled solve in the multigroup driver
  lb::BatchTable batch;             // independent problems (set by lb_op_set_problems)
  std::map<int, lb::SweepTune> btune;  // autotuner state of batched sweeps, per group-batch size
  std::map<int, lb::SweepTune> btune_pos;  // the same for batched one-group sweeps in sweep order, by lb::kOrder*
  std::vector<int32_t> h_nbr;       // int32 [NE][4] neighbours (-1 boundary)
  std::vector<double> h_detJ;       // [NE] 6 x volume
  std::vector<double> h_vert;       // [NE][4][3] vertex coordinates in local node order
  lb::DevBuf d_vert;                // the same on the device (uploaded on the first beam)
  lb::DevBuf w_beam;                // beam scratch: attenuation, failure counter, moment coefficients
  lb::DevBuf w_phi0;                // projected uncollided fluence of the host beam entry points
  // learned right preconditioner (per cell-node MLP), see lb_mlp.hip
  bool mlp = false;
  int mlp_fmt = 1;                  // LB_FMT_BF16 (MFMA) or LB_FMT_FP32
  double mlp_min_tau = 0.0;         // groups optically thinner than this solve without the network
  int moment_fmt = 4;               // LB_FMT_FP64, or the MFMA format of D psi in the Arnoldi steps
  lb::DevBuf w_mix;                 // reduced-precision D psi: packed psi and FP32 moments
  lb::DevBuf d_mlp_w;               // FP32 parameters, natural layout
  lb::DevBuf d_mlp_pack;            // BF16 weights in MFMA lane order
  lb::DevBuf d_elem_h;              // [NE] element size (6 V)^(1/3)
  lb::DevBuf w_feat;                // training/feature scratch
  std::vector<lb::SweepGraph> graphs;
  // runtime
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_sweep_ms = 0.f;
  // workspaces
  lb::DevBuf w_a, w_a2, w_b, w_c, w_phi, w_red, w_krylov, w_host, w_mg, w_acc, w_S;
  lb::DevBuf w_arn;                 // device-resident Arnoldi scalars (lb::ArnoldiDev)
  lb::DevBuf w_tq;                  // T^-1 q of the current solve, for its true residual
  lb::DevBuf w_tail;                // batched residual checks of the multigroup solve (freed after it)
  lb::DevBuf w_so;                  // one group's angular flux in sweep order (within-group solves)
  lb::DevBuf w_so32;                // the same in FP32 (LB_FLUX_SWEEP_ORDER_FP32)
};

namespace lb {

// lb_setup.cpp
int upload(lb_op* op, DevBuf& dst, const void* src, size_t bytes);
int compute_geometry(lb_op* op, const double* rg, const int64_t* node, const int64_t* nbr, const int64_t* mat,
                     int base);
int compute_schedule(lb_op* op, const int64_t* nbr_c);
int build_angular(lb_op* op);
// Real orthonormal spherical harmonics Y_lm(k), l <= L, at index l^2 + l + m.
void real_sph_harm(int L, const double* k, double* y);

// lb_trace_plan.cpp
// Ring and staging slot counts of the LDS sweep for G groups per cell.
int trace_plan_sizes(int lds_bytes, int G, int* ring, int* stage);
// Plans every direction of a level-ordered task array (tasks [NI][NE], level
// offsets [NI][maxlev + 1]); uslot [NI][NE][4], wslot [NI][NE], stage_off
// [NI][maxlev + 2] (entry indices into entries, 2 ints per entry); stats: upstream
// traces from the ring, staged, one level back, staging overflow; gprod
// (optional) [NI][NE]: 1 where a consumer reads the cell's flux from global
// memory (a staging copy or a direct read).
int build_trace_plan(int64_t NI, int64_t NE, int maxlev, const int32_t* nlev, const int32_t* lvl_off,
                     const uint32_t* tasks, const int32_t* nbr, int ring, int stage, int32_t* uslot, int32_t* wslot,
                     int32_t* stage_off, std::vector<int32_t>& entries, int64_t stats[4],
                     uint8_t* gprod = nullptr);
LevelTable* level_table(lb_op* op, int G, int* rc);

// lb_kernels.hip (launchers; all enqueue on op->stream)
int launch_sweep(lb_op* op, const double* q, double* psi, int g0, int G);
// One group's sweep with the flux written in sweep order: psi_so[i][pos] holds the
// flux of the cell at position pos of direction i's device sweep order.  The
// source q is in the natural layout (then it must not alias psi_so) or, with
// source_in_order, in sweep order too (in place allowed).
int launch_sweep_pos(lb_op* op, const double* q, double* psi_so, int g0, bool source_in_order = false);
// One group's sweep (the LDS kernel) writing the flux in sweep order as FP32 to
// psi32_so; psi_so (FP64, sweep order) receives only the cells whose flux a later
// cell reads from global memory, and holds nothing else of use.
int launch_sweep_f32(lb_op* op, const double* q, double* psi_so, float* psi32_so, int g0);
// phi = D psi for one group's FP32 sweep-order flux (FP64 sums).
int launch_moments_pos32(lb_op* op, const float* psi32_so, double* phi);
// phi = D psi_so for one group's sweep-order flux (bit-identical to launch_moments
// on the same flux in natural layout).
int launch_moments_pos(lb_op* op, const double* psi_so, double* phi);
// One group's M S phi (S = sigma_l(mat, g0) when scaled) written in sweep order.
int launch_moments_to_order(lb_op* op, const double* phi, double* psi_so, int g0, bool scaled = true);
// Sweep-order flux of one group to the natural layout.
int launch_flux_to_natural(lb_op* op, const double* psi_so, double* psi);
// Whether the within-group solves keep their angular flux in sweep order (the
// operator's flux order, FP64 moment products of at most 16 moments; with a
// problem set, only when its sweeps are the persistent kernel's, whose
// problem-major order the positions then follow).
bool sweep_order_flux(const lb_op* op);
// Whether they keep it in sweep order as FP32 (LB_FLUX_SWEEP_ORDER_FP32; single
// problems, FP64 moment products of at most 16 moments).
bool sweep_order_fp32(const lb_op* op);
// One sweep with the direction-resident kernel at a fixed workgroup size (no autotuning).
int launch_sweep_direction(lb_op* op, const double* q, double* psi, int g0, int G, int block);
int launch_moments(lb_op* op, const double* psi, double* phi, int G, int64_t ncols);
int launch_moments_to_discrete(lb_op* op, const double* phi, double* psi, int g0, int G, int accumulate,
                               bool scaled = true);
int launch_downscatter(lb_op* op, const double* phi_all, const double* S, int g, int G, const double* qmom,
                       double* acc);
int launch_scatter_dense(lb_op* op, const double* x, double* y, int g0, int G);
int launch_axpby(hipStream_t s, int64_t n, double a, const double* x, double b, const double* y, double* z);
int launch_scale(hipStream_t s, int64_t n, double a, double* x);
int launch_delta_down(lb_op* op, int64_t n, double* x);
void release_blas(lb_op* op);
int launch_copy(hipStream_t s, int64_t n, const double* x, double* y);
// Blocks of the deterministic dot product's first stage (fixed grid).
constexpr int kDotPartials = 2048;
// Enqueues res[0] = <x, y> through part[kDotPartials] (fixed grid and reduction
// tree, so the bits do not vary from run to run); no host synchronisation.
int launch_dot(hipStream_t s, int64_t n, const double* x, const double* y, double* part, double* res);
int dot(lb_op* op, int64_t n, const double* x, const double* y, double* out);
// out[i] = <w, V_i> for i < c (V_i = V + i ld) and, with self, out[c] = <w, w>;
// one host synchronisation.
int multi_dot(lb_op* op, int64_t n, const double* V, int64_t ld, int c, const double* w, bool self, double* out);
// w <- (w - sum_{i<c} h_i V_i) * scale
int multi_axpy(lb_op* op, int64_t n, const double* V, int64_t ld, int c, const double* h, double scale, double* w);

// Device-resident scalars of one restarted-GMRES cycle: the Hessenberg matrix
// H ((m + 1) x m, column-major), the Givens rotations (cs, sn), the rotated
// right-hand side g, and the Gram-Schmidt projections of the current step.  An
// Arnoldi step runs its Gram-Schmidt passes, the decision on the second pass, the
// Hessenberg column, the rotations and the residual estimate on the device, and
// the host reads the step's outcome (ArnoldiCtl) once, after the step.
enum { kArnPass2 = 0, kArnDone1 = 1, kArnDone2 = 2, kArnInvariant = 3, kArnNaN = 4 };
struct ArnoldiCtl {
  double scale;   // factor of the pending orthogonalisation update
  double nu;      // norm of the orthogonalised vector (0: invariant subspace, NaN)
  double resn;    // residual estimate |g_{k+1}| after the step
  double pad;
  int state;      // kArn*
  int pad2[7];
};
struct ArnoldiDev {
  int m = 0;
  ArnoldiCtl* ctl = nullptr;
  double *H = nullptr, *cs = nullptr, *sn = nullptr, *g = nullptr, *hk = nullptr, *d = nullptr, *coef = nullptr,
         *y = nullptr;
};
int arnoldi_reserve(lb_op* op, int m, ArnoldiDev* a);
// g <- (beta, 0, ..., 0): the start of a cycle.
int arnoldi_start(lb_op* op, const ArnoldiDev& a, double beta);
// Orthonormalises V_{k+1} (stride n) against V_0 .. V_k (classical Gram-Schmidt,
// once more when the first pass removes more than half of its squared norm),
// writes column k of H (of I - K when shifted), applies the rotations and
// updates g; one host synchronisation, whose result lands in *out.
int arnoldi_step(lb_op* op, int64_t n, const double* V, int k, const ArnoldiDev& a, bool shifted, ArnoldiCtl* out);
// sol <- sol + V y, y = R^-1 g over the first k columns; no host synchronisation.
int arnoldi_update(lb_op* op, int64_t n, const double* V, int k, const ArnoldiDev& a, double* sol);
int launch_fluence(lb_op* op, const double* v, double* phi);
// lb_mfma.hip: phi [nmom][NENK] = D psi of one group with the operator's reduced
// input format (moment_fmt), FP32 accumulation, result widened to FP64.
int launch_moments_reduced(lb_op* op, const double* psi, double* phi);

// lb_groups.hip (group-batched multigroup solve)
int launch_batch_source(lb_op* op, const double* phi_all, const double* S, const double* qmom, int g0, int B,
                        double* acc);
int launch_batch_transfer(lb_op* op, const double* S, int g0, int B, const double* v, double* out);
int launch_batch_store(lb_op* op, const double* Phi, int g0, int B, double* phi_all);
int launch_batch_gather(lb_op* op, const double* phi_all, int g0, int B, double* Phi);
int launch_group_scale(lb_op* op, int64_t n, int B, const double* w, double* x);
int group_dot(lb_op* op, int64_t n, int B, const double* x, const double* y, double* out);

// lb_mlp.hip
int mlp_set(lb_op* op, const float* params, int nparams, int fmt);
int mlp_apply(lb_op* op, const double* r, double* z, int g0);
int mlp_apply_noalloc(lb_op* op, const double* r, double* z, int g0);
int mlp_features(lb_op* op, const double* r, int g0, float* feat, double* scale);

// lb_beam.hip
int beam_uncollided(lb_op* op, const lb_beam_desc* beam, const double* attn, int64_t e0, int64_t e1, double* d_phi0);
int beam_moments(lb_op* op, const lb_beam_desc* beam, const double* d_phi0, double* d_fcs, double* d_unc);
int beam_angular_fluence(lb_op* op, const lb_beam_desc* beam, const double* d_phi0, double* d_unc);
// d_s [NI][NENK] += the first-collision source of group g0 (FCSSource.f90:134-215)
int beam_source(lb_op* op, const lb_beam_desc* beam, int g0, double* d_s);
// d_unc [NI][NENK] += the uncollided angular fluence for attenuation attn[Nm]
int beam_add_angular_fluence(lb_op* op, const lb_beam_desc* beam, const double* attn, double* d_unc);

// lb_batch.hip
int set_problems(lb_op* op, int P, const int64_t* eoff);
SegTable* seg_table(lb_op* op, int rows, int* rc);
int seg_dot(lb_op* op, SegTable* t, const double* x, const double* y, double* out);
int seg_axpy(lb_op* op, SegTable* t, const double* coef, const double* x, double* y);
int seg_scale(lb_op* op, SegTable* t, const double* coef, double* x);
int batch_solve_dev(lb_op* op, const double* qmom, const lb_solver_opts* opts, double* phi, lb_solve_stats* stats);

// lb_krylov.cpp
int boltzmann_matvec_dev(lb_op* op, double* x, int g0);
int boltzmann_matinv_dev(lb_op* op, double* s, int g0, const lb_solver_opts* opts, lb_solve_stats* st);
int scatter_dev(lb_op* op, const double* x, double* y, int g0, int G);
int multigroup_solve_dev(lb_op* op, const double* qmom, const double* transfer, const lb_solver_opts* opts,
                        double* phi_out, lb_solve_stats* stats);
int moment_matvec_dev(lb_op* op, double* v, int g0);
int moment_rhs_dev(lb_op* op, const double* q, double* b, int g0);
int gmres_snapshots_dev(lb_op* op, const double* q, const lb_solver_opts* o, double* phi, double* snaps,
                        int max_snaps, int* nsnaps, lb_solve_stats* st);
// Optical thickness of group g: the largest sigma_t over the materials times the
// cube root of the mesh volume.
double group_thickness(const lb_op* op, int g);

}  // namespace lb
