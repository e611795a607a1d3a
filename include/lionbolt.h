/*
 * lionbolt.h - C ABI of the MI355X transport-solve library (liblbhip.so).
 *
 * The entry points replace the type-bound procedures of the reference's
 * abstract LinearOperator stack (src/operators/LinearOperatorClass.f90:43-112,
 * re-exported by src/LionboltAPI.f90:1-11).  A Fortran caller binds them with
 * iso_c_binding (lionbolt_amd/fortran/LionboltHIP.f90); other callers use the
 * same symbols through ctypes or plain C.
 *
 * Conventions
 *  - Host pointers, FP64, synchronous on return, unless the name ends in _dev.
 *  - A space-angle vector uses the reference layout v[sdof + i*NENK], 0-based,
 *    sdof = 4*e + k (LinearOperatorClass.f90:30-41,140,166-172).
 *  - _dev entry points take device pointers in the group-batched layout
 *    v[((i*NE + e)*4 + k)*G + g] (for G = 1 this is the reference layout) and
 *    run on the operator's HIP stream; they return before the work completes
 *    unless stated otherwise.
 *  - Every entry point returns 0 on success or a negative LB_E* code, and
 *    lb_last_error() describes the last failure of the calling thread.  The
 *    reference stops the process instead (stophere); the Fortran shim keeps
 *    that behaviour by calling stophere(lb_last_error()) on a non-zero code.
 *  - One handle per operator; a handle is not thread-safe.
 */
#ifndef LIONBOLT_H
#define LIONBOLT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LB_OK = 0,
  LB_E_ARG = -1,        /* invalid argument or size mismatch */
  LB_E_STAGNANT = -2,   /* SweepOrder.f90:178-184 "Sweep is stagnant" */
  LB_E_NOT_READY = -3,  /* operator used before Build (BoltzmannOperator.f90:73-74) */
  LB_E_HIP = -4,        /* HIP runtime failure */
  LB_E_NOMEM = -5,      /* device or host allocation failed */
  LB_E_NAN = -6,        /* "RESIDUAL NORM is NaN." (SPARSKIT.f90:123-125) */
  LB_E_MAXIT = -7,      /* GMRES ipar(1) = -1 (SPARSKIT.f90:159-162) */
  LB_E_BREAKDOWN = -8   /* GMRES ipar(1) = -3 (SPARSKIT.f90:166-169) */
};

enum { LB_SOLVER_GMRES = 0, LB_SOLVER_SI = 1, LB_SOLVER_FGMRES = 2 };
enum { LB_SPACE_PSI = 0, LB_SPACE_MOMENT = 1 };
enum { LB_SCATTER_MOMENT = 0, LB_SCATTER_DENSE = 1 };
enum {
  LB_SWEEP_LEVELS = 0,
  LB_SWEEP_GRAPH = 1,
  LB_SWEEP_DIRECTION = 2,
  LB_SWEEP_AUTO = 3,
  LB_SWEEP_PERSISTENT = 4,
  LB_SWEEP_DIRECTION_LDS = 5
};

typedef struct lb_op lb_op;

/* MeshClass fields the operators read (Geometry.f90:25-69).  Arrays are
 * borrowed for the duration of lb_op_create only. */
typedef struct {
  int64_t NE;            /* 4-node tetrahedra */
  int64_t NK;            /* global nodes */
  const double* rg;      /* rg(3,NK): x,y,z of each node, node-major */
  const int64_t* node;   /* node(4,NE): global nodes of each element */
  const int64_t* nbr;    /* neighbors(4,NE) across local faces [1,2,3],[1,2,4],[1,3,4],[2,3,4] */
  const int64_t* mat;    /* el%mat: material of each element */
  int index_base;        /* 1: Fortran indices, boundary neighbour 0; 0: C indices, boundary -1 */
} lb_mesh_desc;

/* Most elements of one operator: sweep records keep a cell index in 24 bits. */
#define LB_MAX_ELEMENTS 16777215

/* Most materials of one operator (material indices 1..256): sweep records keep
 * the material in 8 bits. */
#define LB_MAX_MATERIALS 256

/* Highest scattering order: PNSCATTERING 15, the highest the reference's
 * input decks use (validation/electrons_incident/Au_1MeV_0deg/lb.in and its siblings). */
#define LB_MAX_ORDER 15

/* AngularClass fields the operators read (AngularSpace.f90:5-20). */
typedef struct {
  int64_t NI;            /* directions */
  const double* k;       /* k(3,NI) unit propagation directions */
  const double* w;       /* w(NI) weights (sum 4 pi) */
  int L;                 /* scattering order: YY(:,:,0:L), 0 <= L <= LB_MAX_ORDER */
} lb_quad_desc;

typedef struct {
  int kind;              /* LB_SOLVER_GMRES (reference default), LB_SOLVER_SI, LB_SOLVER_FGMRES
                            (flux moments, learned right preconditioner, lb_op_set_preconditioner) */
  int space;             /* LB_SPACE_PSI: Krylov on psi as the reference; LB_SPACE_MOMENT: on flux moments */
  int restart;           /* GMRES(m); the reference hard-codes 10 (BoltzmannOperator.f90:135) */
  int maxit;             /* <= 0: no limit (SPARSKIT.f90:89) */
  double rtol, atol;     /* 1e-8, 1e-12 (SPARSKIT.f90:100-101, SourceIteration.f90:14-15) */
  int verbose;           /* 1: print the reference's residual table to stdout */
} lb_solver_opts;

typedef struct {
  int status;            /* 0 converged, else a negative LB_E* code */
  int iterates;          /* the reference's "Total number of iterates" */
  int matvecs;           /* operator applications inside the solver */
  int sweeps;            /* transport sweeps including the T^-1 q of MatInv */
  double final_norm;     /* last printed residual norm (Arnoldi estimate / SI increment) */
  double true_residual;  /* ||T^-1 q - (I - T^-1 K) psi|| / ||T^-1 q||, recomputed in FP64 */
  double seconds;        /* wall time of the solve */
  int fallback;          /* LB_SOLVER_FGMRES: Arnoldi step after which the learned preconditioner
                            was dropped for stagnation (0: never; -1: not used, the group is below
                            the threshold of lb_op_set_preconditioner_threshold) */
} lb_solve_stats;

/* Op%Build(mesh) + Op%Build(mesh, angular): normals, geometry and the upwind
 * schedule (TransportOperator.f90:172-388, ScatteringOperator.f90:100-213,
 * BoltzmannOperator.f90:171-296). */
int lb_op_create(const lb_mesh_desc* mesh, const lb_quad_desc* quad, lb_op** out);

/* Op%Build(angular, XS): sigma_t(Nm) and Legendre moments sigma_s(0:L, Nm)
 * stored [m][l] (TransportOperator.f90:390-429, ScatteringOperator.f90:215-277). */
int lb_op_set_physics(lb_op* op, int Nm, const double* sigma_t, int L, const double* sigma_s);

/* Multigroup physics for the batched device entry points: sigma_t[g][m],
 * within-group sigma_s[g][m][l]. */
int lb_op_set_physics_groups(lb_op* op, int G, int Nm, const double* sigma_t, int L, const double* sigma_s);

/* Scattering kernel choice: LB_SCATTER_DENSE applies YYS(NI,NI) as the
 * reference does (ScatteringAlgebra.f90:30-53); LB_SCATTER_MOMENT applies the
 * identical K = M diag(sigma_l) D (default). */
int lb_op_set_scatter_mode(lb_op* op, int mode);

/* How a sweep is issued: LB_SWEEP_LEVELS launches one kernel per global level
 * on the operator's stream; LB_SWEEP_GRAPH captures those launches once per
 * (source, flux, group range) into a HIP graph and replays it;
 * LB_SWEEP_DIRECTION runs one kernel in which each workgroup owns one
 * direction and walks its levels with workgroup barriers;
 * LB_SWEEP_DIRECTION_LDS does the same and also keeps upstream traces in LDS
 * slots fixed by a slot plan built on first use per group-batch size (ring
 * slots by Belady's rule, staging copies one level ahead, global memory for
 * the rest; see lb_trace_plan_host); LB_SWEEP_AUTO (default) times the graph
 * mode, the direction-resident mode at 256, 512 and 1024 threads and, below 8
 * groups per cell, the LDS one at 256 and 512 threads on the first sweeps of
 * each group-batch size (two sweeps each, synchronising after each) and then
 * keeps the fastest.  All modes give bit-identical results.  No reference counterpart (the reference sweeps one
 * direction per OpenMP thread). */
int lb_op_set_sweep_mode(lb_op* op, int mode);

/* The launch configuration LB_SWEEP_AUTO settled on for a group batch of G:
 * *mode is LB_SWEEP_GRAPH, LB_SWEEP_DIRECTION, LB_SWEEP_DIRECTION_LDS or LB_SWEEP_PERSISTENT (with
 * *threads its workgroup size and *per_cu the persistent grid's workgroups per
 * compute unit), or LB_SWEEP_AUTO while that batch size is still being timed. */
int lb_op_sweep_choice(lb_op* op, int G, int* mode, int* threads, int* per_cu);

/* Workgroups per compute unit of the persistent sweep's grid (1..8, with
 * threads x per_cu <= 2048); the workgroup size is lb_op_set_direction_block's. */
int lb_op_set_workgroups_per_cu(lb_op* op, int per_cu);

/* Batches of independent problems (no reference counterpart: the reference
 * solves one problem per run).  The operator's mesh is the disjoint union of P
 * problems in consecutive element ranges [elem_off[p], elem_off[p+1]) (0-based,
 * elem_off[0] = 0, elem_off[P] = NE) with no face between ranges; give each
 * problem its own material to give it its own cross sections.  Enables
 * LB_SWEEP_PERSISTENT: one launch of a fixed grid that pulls (problem,
 * direction) items from a work queue, largest problems first, replayed from a
 * HIP graph; LB_SWEEP_AUTO then times it (several workgroup sizes and grid
 * occupancies) against the level launches. */
int lb_op_set_problems(lb_op* op, int P, const int64_t* elem_off);

/* ---- learned right preconditioner (no reference counterpart) -------------
 * A per-cell-node MLP 18 -> 64 -> 64 -> 64 -> 16 (tanh hidden layers): inputs the
 * 16 flux moments of a residual at one spatial dof (divided by the residual's
 * root-mean-square s), log(sigma_t h) and sigma_s0/sigma_t of the dof's element
 * (h = (6 V)^(1/3)); outputs the error's 16 moments there (times s).  Used as
 * the right preconditioner z = P(v) of flexible GMRES(m) on flux moments
 * (LB_SOLVER_FGMRES); when the residual estimate falls by less than a factor
 * 8 over the last three Arnoldi steps, or P returns a non-finite or zero
 * vector, the solve continues as plain GMRES.  Parameters: W1[18][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
 * W4[64][16] b4[16], FP32, W[i][j] from input i to output j. */
#define LB_MLP_PARAMS 10576

/* Sets the network; fmt LB_FMT_BF16 (MFMA inference) or LB_FMT_FP32. */
int lb_op_set_preconditioner(lb_op* op, const float* params, int nparams, int fmt);

/* The a-priori choice of LB_SOLVER_FGMRES: a group whose optical thickness
 * (largest sigma_t over the materials times the cube root of the mesh volume)
 * is below min_thickness is solved without the network (plain flux-moment GMRES;
 * lb_solve_stats.fallback = -1).  Default 0 (every group uses it); +inf turns the
 * network off.  The caller derives the threshold from the training problems. */
int lb_op_set_preconditioner_threshold(lb_op* op, double min_thickness);

/* Input precision of the discrete-to-moment product D psi inside the Arnoldi
 * steps of the flux-moment GMRES (K = D T^-1 M S): LB_FMT_FP64 (default, the
 * FP64 kernels) or an MFMA format of lb_gemm_d_dev (psi packed column-major in
 * the format, FP32 accumulation, the moments widened to FP64).  The right-hand
 * side, psi and the true residual stay FP64.  Reduced formats need 256
 * directions and P3 (16 moments).  No reference counterpart (the reference
 * computes K in FP64). */
int lb_op_set_moment_precision(lb_op* op, int fmt);

/* d_z [16][NENK] = P(d_r [16][NENK]) for group g0 (single-group moment vectors). */
int lb_mlp_apply_dev(lb_op* op, const double* d_r, double* d_z, int g0);

/* lb_mlp_apply_dev without any allocation or host synchronisation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight solver loops).  The residual's norm
 * is reduced in a file-scope device array of LB_NOALLOC_SCRATCH_BYTES per GPU
 * (its need does not grow with the mesh, so no size is refused for it), held
 * until the process ends; the network must be set first (lb_op_set_preconditioner,
 * which also uploads the element sizes the inputs use).  Calls on one GPU share
 * that array: calls that overlap on one GPU race.  Asynchronous on the
 * operator's stream; the same output as lb_mlp_apply_dev. */
#define LB_NOALLOC_SCRATCH_BYTES 16448
int lb_mlp_apply_noalloc_dev(lb_op* op, const double* d_r, double* d_z, int g0);

/* The network's inputs for d_r: d_feat [NENK][18] (FP32) and the scale s. */
int lb_mlp_features_dev(lb_op* op, const double* d_r, int g0, float* d_feat, double* scale);

/* Trains the network in FP32 by minibatch Adam (batch <= 4096 rows drawn with
 * SplitMix64 stream `seed`) on device rows d_X [N][18] -> d_Y [N][16] (mean
 * squared error); params [LB_MLP_PARAMS] in (initial) and out (trained) on the
 * host; loss_hist [steps] (host, may be NULL). */
int lb_mlp_train_dev(const float* d_X, const float* d_Y, int64_t N, float* params, int nparams, int steps,
                     int batch, float lr, uint64_t seed, float* loss_hist);

/* Gradient of the mean squared error over rows 0..B-1 (B <= 4096) at params (host
 * in), grad [LB_MLP_PARAMS] (host out), without an update. */
int lb_mlp_grad_dev(const float* d_X, const float* d_Y, int B, const float* params, int nparams, float* grad,
                    float* loss);

/* Within-group flux-moment system (I - D T^-1 M S) phi = D T^-1 q of group g0
 * (single-group moment vectors [nmom][NENK]): d_b = D T^-1 d_q (d_q: discrete
 * source, NENK*NI), and d_y = d_x - D T^-1 M S d_x. */
int lb_moment_rhs_dev(lb_op* op, const double* d_q, double* d_b, int g0);
/* Discrete source of single-group source moments: d_q [NI][NENK] = M d_qmom (no
 * cross sections applied). */
int lb_discrete_source_dev(lb_op* op, const double* d_qmom, double* d_q);
int lb_moment_matvec_dev(lb_op* op, const double* d_x, double* d_y, int g0);

/* Flux-moment GMRES(m) of group g0 from x0 = b / sigma_t(1) that records the
 * iterate x_k after every Arnoldi step (d_snaps [max_snaps][nmom*NENK]) and
 * returns the solution in d_phi: residual/error pairs for training the
 * preconditioner come from the solver itself. */
int lb_gmres_snapshots_dev(lb_op* op, const double* d_q, const lb_solver_opts* opts, double* d_phi,
                           double* d_snaps, int max_snaps, int* nsnaps, lb_solve_stats* stats);

/* Solves every problem of the batch, single group: flux-moment GMRES(m) per
 * problem, all problems in lockstep with one sweep launch per Arnoldi step
 * (converged problems leave the sweep).  d_qmom [nmom][NENK]: moments of each
 * problem's external source (discrete source = M qmom); d_phi [nmom][NENK]:
 * collided flux moments out.  stats[P]: per-problem status, Arnoldi steps and
 * true transport residual.  opts: restart, maxit, rtol, atol (kind and space
 * are ignored). */
int lb_batch_solve_dev(lb_op* op, const double* d_qmom, const lb_solver_opts* opts, double* d_phi,
                       lb_solve_stats* stats);

/* ---- planar external beam (the reference's ExternalBeam / FieldGeometry,
 * src/sources/Sources.f90:13-80) -------------------------------------------- */
enum { LB_CUTOUT_NONE = 0, LB_CUTOUT_CIRCLE = 1, LB_CUTOUT_RECTANGLE = 2 };

typedef struct {
  double axis[3];    /* FieldGeometry%axis: unit direction of the beam's particles */
  double origin[3];  /* FieldGeometry%origin: a point of the source plane, upstream of the mesh */
  double e1[3];      /* FieldGeometry%e1, %e2: orthonormal beam frame normal to the axis (the */
  double e2[3];      /* rectangle cutout is measured along them) */
  int cutout;        /* FieldGeometry%cutout: LB_CUTOUT_* */
  double params[4];  /* FieldGeometry%cutoutparams: circle radius params[1]; rectangle widths params[2], params[3] */
  double n;          /* ExternalBeam%n: particles in the beam */
} lb_beam_desc;

/* The beam's uncollided fluence projected on the nodal basis, the phi0ek that
 * FCSSource.f90:134-176 and AddAngularFluence.f90:109-151 build:
 * d_phi0 [NENK] = IMr (sum_kq w_kq u_k(kq) phi0q(kq)) per element, phi0q =
 * n exp(-sum_m attn[m] l_m) at the four tetrahedral quadrature points that lie in
 * the field (InField.f90), l_m the path length in material m from the point back
 * along the axis to the mesh surface (RayTracing.f90:149-199,
 * FluenceQuadrature.f90:3-33; traced element to element on the device).
 * attn [Nm] on the host: sigma_t of each material.  Synchronous. */
int lb_beam_uncollided_dev(lb_op* op, const lb_beam_desc* beam, const double* attn, double* d_phi0);

/* lb_beam_uncollided_dev on the elements [e0, e1) only (d_phi0 entries 4 e0 ..
 * 4 e1 - 1 written).  No reference counterpart (one beam lights the whole mesh
 * there): a batch of independent problems in consecutive element ranges, each
 * lit by its own beam (lb_op_set_problems). */
int lb_beam_uncollided_range_dev(lb_op* op, const lb_beam_desc* beam, const double* attn, int64_t e0, int64_t e1,
                                 double* d_phi0);

/* Flux moments of a beam of projected fluence d_phi0 [NENK] (either output may be
 * NULL): d_fcs [nmom][NENK] = Y_lm(axis) phi0, whose discrete form with
 * sigma_l(mat) applied, M diag(sigma_l) d_fcs, is the first-collision source of
 * FCSSource.f90:178-215 (addition theorem); d_unc [nmom][NENK] = (D g)_lm phi0,
 * the moments of the uncollided angular fluence phi0 g_i of
 * AddAngularFluence.f90:153-166, g_i = sum_l (2l+1)/(4 pi) P_l(k_i . axis).
 * Synchronous. */
int lb_beam_moments_dev(lb_op* op, const lb_beam_desc* beam, const double* d_phi0, double* d_fcs, double* d_unc);

/* AddAngularFluence.f90:153-166: d_unc [NI][NENK] += phi0 g_i.  Synchronous. */
int lb_beam_angular_fluence_dev(lb_op* op, const lb_beam_desc* beam, const double* d_phi0, double* d_unc);

/* ExternalBeam%Source => FCSSource (FCSSource.f90:3-215, planar beam) on a
 * host vector in the reference layout s[sdof + i NENK] (n = NENK NI): s += the
 * first-collision source, phi0 attenuated by the operator's sigma_t (the
 * single-group physics, as FluenceQuadrature(OpK%XS%t)) and weighted by
 * sum_l (2l+1)/(4 pi) P_l(k_i . axis) sigma_l(material).  Synchronous. */
int lb_beam_source(lb_op* op, const lb_beam_desc* beam, double* s, int64_t n);

/* ExternalBeam%AddAngularFluence (AddAngularFluence.f90:3-166, planar) on a
 * host vector: unc += phi0 g_i, phi0 attenuated by attn[material].
 * Synchronous. */
int lb_beam_add_angular_fluence(lb_op* op, const lb_beam_desc* beam, const double* attn, double* unc, int64_t n);

/* Workgroup size of the direction-resident sweeps: 256, 512, 768 or 1024
 * (default); 768 is not a persistent-sweep size. */
int lb_op_set_direction_block(lb_op* op, int threads);

/* TransportOp%MatInv (TransportOperator.f90:72-151): s <- T^-1 s in place. */
int lb_transport_matinv(lb_op* op, double* s, int64_t n);

/* ScatteringOp%MatVec (ScatteringOperator.f90:39-70): x <- K x in place.
 * After lb_op_build_delta_down, K is the delta-down kernel. */
int lb_scatter_matvec(lb_op* op, double* x, int64_t n);

/* ScatteringOp%BuildDeltaDown (ScatteringOperator.f90:291-319): K becomes
 * XSdd(mat) times the identity in angle (ApplyDeltaDown,
 * ScatteringAlgebra.f90:62-90), xsdd[m] for each of the Nm mesh materials
 * (Nm >= the mesh's material count).  The reference's energy-coupling loop
 * applies it to a source group's angular flux when the cross sections carry a
 * delta-down row (Main.f90:610-624).  Only lb_scatter_matvec applies it: the
 * Boltzmann and transport entry points of a delta-down handle return
 * LB_E_ARG until lb_op_set_physics, which restores the Legendre kernel as
 * ScatteringOp%Physics does (ScatteringOperator.f90:231). */
int lb_op_build_delta_down(lb_op* op, int Nm, const double* xsdd);

/* BoltzmannOp%MatVec (BoltzmannOperator.f90:59-104): x <- x - T^-1 K x. */
int lb_boltzmann_matvec(lb_op* op, double* x, int64_t n);

/* BoltzmannOp%MatInv (BoltzmannOperator.f90:106-151): s (source q) <- psi. */
int lb_boltzmann_matinv(lb_op* op, double* s, int64_t n, const lb_solver_opts* opts, lb_solve_stats* stats);

/* SpaceAngleVector%Fluence (LinearOperatorClass.f90:154-174): phi(NENK) = sum_i w_i v_i. */
int lb_fluence(lb_op* op, const double* v, double* phi);

/* Schedule of SweepOrder.f90:105-207 for parity checks: sl[i][es] (0-based
 * element visited at position es for direction i) and nlev[i]. */
int lb_op_get_schedule(lb_op* op, int32_t* sl, int32_t* nlev);

/* The host half of lb_op_create without a device: the schedule sl[i][es],
 * nlev[i] and normals n[e][f][3] (any output may be NULL).  Lets the ordering
 * be checked where no GPU is present. */
int lb_schedule_host(const lb_mesh_desc* mesh, const lb_quad_desc* quad, int32_t* sl, int32_t* nlev, double* normals);

/* The LDS slot plan of the direction-resident LDS sweep (LB_SWEEP_DIRECTION_LDS)
 * for G groups per cell and lds_bytes of LDS, built on the host for the
 * schedule lb_schedule_host returns; no device calls.  No counterpart in the
 * reference, which has no LDS: it lets the plan be checked where no GPU is
 * present.  Outputs (NULL skips one):
 *   uslot     [NI][NE][4]  LDS slot of each upstream trace, -1 = global memory
 *                          (also -1 on faces that are not upstream)
 *   wslot     [NI][NE]     slot each cell writes, -1 = none
 *   stage_off [NI][maxlev + 2]  first staging entry of each consumer level
 *   stage     2 per entry: (LDS slot, element copied from global memory), filled
 *             only if *nstage holds at least the entry count; *nstage returns it
 *   sizes     {ring slots, staging slots per level parity}
 *   stats     upstream traces from the ring, staged, one level back, overflow
 *   gprod     [NI][NE]     1 = some consumer reads the cell's flux from global
 *                          memory (a staging copy, or a consumer with any trace
 *                          outside LDS, which reads all its traces there): the
 *                          cells whose FP64 flux the FP32 flux mode stores
 * Any output may be NULL. */
int lb_trace_plan_host(const lb_mesh_desc* mesh, const lb_quad_desc* quad, int G, int lds_bytes, int32_t* uslot,
                       int32_t* wslot, int32_t* stage_off, int32_t* stage, int64_t* nstage, int32_t* sizes,
                       int64_t* stats, uint8_t* gprod);

/* Outward unit normals n[e][f][3] as PostProcess.f90:48-102 computes them. */
int lb_op_get_normals(lb_op* op, double* n);

/* Sizes: NE, NI, number of moments (L+1)^2, max level count. */
int lb_op_info(lb_op* op, int64_t* NE, int64_t* NI, int64_t* nmom, int64_t* maxlev);

/* ---- device-resident entry points (no reference counterpart: they expose
 * the same operators on HBM-resident data so callers skip the PCIe copies) -- */

/* T^-1 over G groups starting at group g0: psi = T^-1 q (in place allowed).
 * Asynchronous: the sweep is enqueued on the operator's stream and may still
 * run on return (lb_op_synchronize waits for it). */
int lb_transport_matinv_dev(lb_op* op, const double* d_q, double* d_psi, int g0, int G);

/* phi[lm][sdof][g] = sum_i D[lm][i] psi[i][sdof][g] (discrete to moment). */
int lb_moments_dev(lb_op* op, const double* d_psi, double* d_phi, int G);

/* psi[i][sdof][g] = sum_lm M[i][lm] sigma_l(mat,g) phi[lm][sdof][g] (+ psi if accumulate). */
int lb_moments_to_discrete_dev(lb_op* op, const double* d_phi, double* d_psi, int g0, int G, int accumulate);

/* Layout of one group's angular flux inside the within-group flux-moment
 * solves (no reference counterpart).  LB_FLUX_NATURAL keeps the operator's
 * layout psi[i][e][k] and sweeps in place.  LB_FLUX_SWEEP_ORDER has the sweep
 * write the flux of direction i at the cell's position in that direction's
 * sweep order, so the cells of one level store to consecutive addresses
 * instead of scattered 32-byte pieces, and D psi gathers it back by element
 * (the default: the 24-group 384k-tet x 256-direction P3 solve 1.18 s against
 * 1.32 s natural, the one-group sweep 3.6 ms against 4.9 ms in place).
 * LB_FLUX_SWEEP_ORDER_SOURCE also has M S phi write the sweep's source in
 * sweep order, and the sweep runs in place on it (3.0 ms), but M S phi's
 * scattered stores then cost more than that saves (1.28 s).  All three give
 * bit-identical solves.  Sweep order applies to FP64 moment products of at
 * most 16 moments; on an operator with a problem set (lb_op_set_problems) it
 * follows the persistent sweep's problem-major order and applies when the
 * sweep mode is LB_SWEEP_AUTO or LB_SWEEP_PERSISTENT (the batched sweep of the
 * 64 uneven beams: 11.0 -> 8.0 ms).  Other operators keep the natural layout
 * whatever is set here.  LB_FLUX_SWEEP_ORDER_FP32 (single problems only; an
 * operator with a problem set runs LB_FLUX_SWEEP_ORDER instead) stores the
 * Arnoldi steps' flux for D psi in FP32: the sweep is always the LDS kernel
 * (at the operator's workgroup size in LB_SWEEP_DIRECTION_LDS, otherwise the
 * fastest of its sizes, timed on the first sweeps), its arithmetic and LDS
 * traces FP64, its FP32 flux the FP64 flux rounded to nearest; the right-hand
 * side T^-1 q, the true residual and the multigroup tail stay FP64.  Not
 * bit-identical to the FP64 layouts (the 24-group solve: 1.07-1.09 s, scalar
 * flux within 6e-9 of the FP64 solve, same iterate counts). */
enum { LB_FLUX_NATURAL = 0, LB_FLUX_SWEEP_ORDER = 1, LB_FLUX_SWEEP_ORDER_SOURCE = 2, LB_FLUX_SWEEP_ORDER_FP32 = 3 };
int lb_op_set_flux_order(lb_op* op, int order);

/* The launch configuration LB_SWEEP_AUTO settled on for one-group sweeps with a
 * sweep-order flux (timed apart from the natural layout's, and per source
 * layout): *mode as in lb_op_sweep_choice, *threads its workgroup size;
 * LB_SWEEP_AUTO while timing. */
int lb_op_sweep_order_choice(lb_op* op, int source_in_order, int* mode, int* threads);

/* One group's T^-1 with the flux in sweep order: d_psi_so[(i NE + pos) 4 + k]
 * is node k of the cell at position pos of direction i's device sweep order.
 * d_q is in the natural layout and must not overlap d_psi_so, or, with
 * source_in_order, in sweep order too, and may then be d_psi_so itself (in
 * place).  Asynchronous like lb_transport_matinv_dev.  No reference
 * counterpart. */
int lb_transport_matinv_order_dev(lb_op* op, const double* d_q, double* d_psi_so, int g0, int source_in_order);

/* psi_so = M S_g0 phi for one group, written in sweep order (the values of
 * lb_moments_to_discrete_dev at the cells' sweep positions). */
int lb_moments_to_discrete_order_dev(lb_op* op, const double* d_phi, double* d_psi_so, int g0);

/* phi = D psi for one group's sweep-order flux (bit-identical to lb_moments_dev
 * on the same flux in the natural layout). */
int lb_moments_order_dev(lb_op* op, const double* d_psi_so, double* d_phi);

/* One group's T^-1 (the LDS sweep) storing the flux in sweep order as FP32 in
 * d_psi32_so: each value is the FP64 flux rounded to nearest FP32, the sweep's
 * arithmetic and its LDS traces staying FP64.  d_psi_so (FP64, sweep order)
 * receives the flux only of the cells that some later cell reads from global
 * memory; its other entries are left as they were.  d_q is natural and must
 * not overlap the outputs.  The solves' LB_FLUX_SWEEP_ORDER_FP32 mode runs the
 * Arnoldi steps' K on it; right-hand side, psi and the true residual stay FP64.
 * No reference counterpart. */
int lb_transport_matinv_order32_dev(lb_op* op, const double* d_q, double* d_psi_so, float* d_psi32_so, int g0);

/* phi = D psi for one group's FP32 sweep-order flux (FP64 sums in the order of
 * lb_moments_order_dev). */
int lb_moments_order32_dev(lb_op* op, const float* d_psi32_so, double* d_phi);

/* One group's sweep-order flux in the natural layout (the two vectors must not
 * overlap). */
int lb_flux_to_natural_dev(lb_op* op, const double* d_psi_so, double* d_psi);

/* Multigroup downscatter solve (the group loop of Main.f90:326-476 with the
 * energy-coupling source of PIEISource, Main.f90:584-629): groups in order,
 * each a within-group BoltzmannOp%MatInv with source
 * M (qmom_g + sum_{g'<g} sigma_l(g'->g) phi_g'), flux moments kept in HBM.
 * d_qmom [G][nmom][NENK] or NULL; transfer (host) [G][G][Nm][L+1] indexed
 * [g'][g]; d_phi [G][nmom][NENK] out; stats[G] out.  Synchronous.  With the
 * flux-moment solvers (LB_SPACE_MOMENT, GMRES or FGMRES) d_phi holds each
 * group's solution phi_g, and psi_g = T^-1 (M S_g phi_g + q_g) and its true
 * transport residual are evaluated after all groups, 8 groups per sweep
 * (each group's stats count those two sweeps).  Otherwise d_phi = D psi_g of
 * each group's BoltzmannOp%MatInv. */
int lb_multigroup_solve_dev(lb_op* op, const double* d_qmom, const double* transfer, const lb_solver_opts* opts,
                            double* d_phi, lb_solve_stats* stats);

/* Energy groups per coupled solve in lb_multigroup_solve_dev (1..16, default 1).
 * With B > 1, batches of B consecutive groups are solved as one flux-moment
 * GMRES system whose operator includes the downscatter inside the batch, so each
 * Arnoldi step is one B-group sweep; batches still follow each other in energy
 * (Gauss-Seidel).  Each group's relative residual meets rtol (the system is
 * scaled per group and solved to rtol / sqrt(B)).  stats[g].sweeps holds the
 * batch's sweeps in its first group and 0 in the others.  No reference
 * counterpart (the reference solves group by group). */
int lb_op_set_group_batch(lb_op* op, int groups);

/* ---- scattering-moment products on MFMA, reduced input precision, FP32
 * accumulation (no reference counterpart: the reference applies the dense
 * FP64 kernel YYS with dgemm, ScatteringAlgebra.f90:30-53).  Formats:
 * LB_FMT_FP32, LB_FMT_BF16, LB_FMT_SPLIT_BF16 (x = hi + lo, three products),
 * LB_FMT_FP16 (power-of-two scale per column).  Reduced operands are stored with
 * the reduced dimension contiguous per column.  stream: a hipStream_t or NULL. */
enum { LB_FMT_FP32 = 0, LB_FMT_BF16 = 1, LB_FMT_SPLIT_BF16 = 2, LB_FMT_FP16 = 3, LB_FMT_FP64 = 4 };

/* src [rows][ncols] FP64 -> dst [ncols][rows] in the format (dst_lo: low part of
 * the split format; inv_scale[ncols]: FP16 column scales 2^-e, written here). */
int lb_pack_dev(int fmt, int rows, int64_t ncols, const double* src, void* dst, void* dst_lo, float* inv_scale,
                void* stream);

/* phi[m][col] (FP32) = sum_i D[m][i] psi[col][i]; D device FP64 [nmom][NI], NI = 256, nmom <= 16. */
int lb_gemm_d_dev(int fmt, int nmom, int NI, const double* D, int64_t ncols, const void* psi, const void* psi_lo,
                  const float* inv_scale, float* phi, void* stream);

/* psi[col][i] (FP32) = sum_m M[i][m] phi[col][m]; M device FP64 [NI][16], NI = 256. */
int lb_gemm_m_dev(int fmt, int NI, int nmom, const double* M, int64_t ncols, const void* phi, const void* phi_lo,
                  const float* inv_scale, float* psi, void* stream);

/* Block until the operator's stream is idle. */
int lb_op_synchronize(lb_op* op);

/* Timing of the most recent device sweep, measured with HIP events on the
 * operator's stream around the level launches: milliseconds. */
int lb_last_sweep_ms(lb_op* op, double* ms);

void lb_op_destroy(lb_op* op);
const char* lb_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LIONBOLT_H */
