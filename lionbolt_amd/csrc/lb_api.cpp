// C ABI of liblbhip: operator lifecycle and the host/device entry points that
// stand in for the reference's LinearOperator procedures.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "lb_internal.hpp"

namespace lb {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }

void clear_graphs(lb_op* op) {
  if (op->stream && !op->graphs.empty()) (void)hipStreamSynchronize(op->stream);  // none may be in flight
  for (auto& sg : op->graphs) (void)hipGraphExecDestroy(sg.exec);
  op->graphs.clear();
}

int DevBuf::reserve(size_t n) {
  if (n <= bytes && p) return LB_OK;
  if (p) {
    hipError_t e = hipFree(p);
    (void)e;
    p = nullptr;
    bytes = 0;
  }
  size_t want = std::max<size_t>(n, 256);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    set_error("device allocation of " + std::to_string(want) + " bytes failed: " + hipGetErrorString(e));
    return LB_E_NOMEM;
  }
  bytes = want;
  return LB_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

}  // namespace lb

using namespace lb;

namespace {

int check_ready(lb_op* op, const char* who) {
  if (!op) {
    set_error(std::string(who) + ": null operator");
    return LB_E_ARG;
  }
  if (op->delta_down) {
    set_error(std::string(who) + " was called on a delta-down scattering kernel, which only "
              "ScatteringOp%MatVec applies (set the physics first).");
    return LB_E_ARG;
  }
  if (!op->phys) {
    set_error(std::string(who) + " was called, but the operator was not ready.");
    return LB_E_NOT_READY;
  }
  return LB_OK;
}

int check_n(lb_op* op, int64_t n, const char* who) {
  if (n != op->NENK * op->NI) {
    set_error(std::string(who) + ": vector length " + std::to_string(n) + " != NENK*NI = " +
              std::to_string(op->NENK * op->NI));
    return LB_E_ARG;
  }
  return LB_OK;
}

int to_device(lb_op* op, const double* h, int64_t n) {
  LB_TRY(op->w_a.reserve(n * sizeof(double)));
  LB_HIP(hipMemcpyAsync(op->w_a.p, h, n * sizeof(double), hipMemcpyHostToDevice, op->stream));
  return LB_OK;
}

int to_host(lb_op* op, double* h, int64_t n) {
  LB_HIP(hipMemcpyAsync(h, op->w_a.p, n * sizeof(double), hipMemcpyDeviceToHost, op->stream));
  LB_HIP(hipStreamSynchronize(op->stream));
  return LB_OK;
}

int upload_physics(lb_op* op) {
  clear_graphs(op);  // captured sweeps hold the physics-table pointer
  LB_TRY(op->d_sigt.reserve(op->h_sigt.size() * sizeof(double)));
  LB_TRY(op->d_sigs.reserve(op->h_sigs.size() * sizeof(double)));
  LB_HIP(hipMemcpy(op->d_sigt.p, op->h_sigt.data(), op->h_sigt.size() * sizeof(double), hipMemcpyHostToDevice));
  LB_HIP(hipMemcpy(op->d_sigs.p, op->h_sigs.data(), op->h_sigs.size() * sizeof(double), hipMemcpyHostToDevice));
  // dense kernel YYS(ip, i, m) = sum_l YY(ip, i, l) XS%s(l, m) for group 0
  // (ScatteringOperator.f90:263-269)
  const int64_t NI = op->NI;
  std::vector<double> yys((size_t)op->Nm * NI * NI, 0.0);
  for (int m = 0; m < op->Nm; ++m)
    for (int l = 0; l <= op->L; ++l) {
      const double sl = op->h_sigs[((size_t)m * op->G + 0) * (op->L + 1) + l];
      const double* yy = &op->h_yy[(size_t)l * NI * NI];
      double* dst = &yys[(size_t)m * NI * NI];
      for (int64_t j = 0; j < NI * NI; ++j) dst[j] += yy[j] * sl;
    }
  LB_TRY(op->d_yys.reserve(yys.size() * sizeof(double)));
  LB_HIP(hipMemcpy(op->d_yys.p, yys.data(), yys.size() * sizeof(double), hipMemcpyHostToDevice));
  op->phys = true;
  op->delta_down = false;   // Physics builds the Legendre kernel (ScatteringOperator.f90:231)
  return LB_OK;
}

}  // namespace

extern "C" {

const char* lb_last_error(void) { return g_error.c_str(); }

int lb_op_create(const lb_mesh_desc* mesh, const lb_quad_desc* quad, lb_op** out) {
  return guarded([&]() -> int {
    if (!mesh || !quad || !out || !mesh->rg || !mesh->node || !mesh->nbr || !mesh->mat || !quad->k || !quad->w) {
      set_error("lb_op_create: null argument");
      return LB_E_ARG;
    }
    if (mesh->NE <= 0 || mesh->NK <= 0 || quad->NI <= 0 || quad->L < 0 || quad->L > LB_MAX_ORDER ||
        (mesh->index_base != 0 && mesh->index_base != 1)) {
      set_error("lb_op_create: invalid sizes (need NE, NK, NI > 0, 0 <= L <= 15, index_base 0 or 1)");
      return LB_E_ARG;
    }
    if (mesh->NE > LB_MAX_ELEMENTS) {
      set_error("lb_op_create: at most 16,777,215 elements per operator (the sweep records hold 24-bit cell "
                "indices); split larger meshes into problems of one batch or several operators");
      return LB_E_ARG;
    }
    for (int64_t e = 0; e < mesh->NE; ++e)
      if (mesh->mat[e] - mesh->index_base >= LB_MAX_MATERIALS) {
        set_error("lb_op_create: at most 256 materials per operator (the sweep records hold 8-bit material "
                  "indices)");
        return LB_E_ARG;
      }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
      (void)hipGetLastError();
      set_error("lb_op_create: no HIP device is visible");
      return LB_E_HIP;
    }
    lb_op* op = new lb_op;
    op->NE = mesh->NE;
    op->NK = mesh->NK;
    op->NI = quad->NI;
    op->NENK = 4 * mesh->NE;
    op->L = quad->L;
    op->h_k.assign(quad->k, quad->k + 3 * quad->NI);
    op->h_w.assign(quad->w, quad->w + quad->NI);
    int rc = LB_OK;
    if (hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&op->ev0) != hipSuccess || hipEventCreate(&op->ev1) != hipSuccess) {
      set_error("lb_op_create: stream or event creation failed");
      rc = LB_E_HIP;
    }
    // a failed host allocation here must not leak the operator
    if (rc == LB_OK) rc = guarded([&] { return compute_geometry(op, mesh->rg, mesh->node, mesh->nbr, mesh->mat,
                                                                mesh->index_base); });
    if (rc == LB_OK) rc = guarded([&] { return build_angular(op); });
    if (rc != LB_OK) {
      lb_op_destroy(op);
      return rc;
    }
    *out = op;
    return LB_OK;
  });
}

int lb_schedule_host(const lb_mesh_desc* mesh, const lb_quad_desc* quad, int32_t* sl, int32_t* nlev, double* normals) {
  return guarded([&]() -> int {
    if (!mesh || !quad || !mesh->rg || !mesh->node || !mesh->nbr || !mesh->mat || !quad->k || !quad->w ||
        mesh->NE <= 0 || mesh->NE > LB_MAX_ELEMENTS || quad->NI <= 0 ||
        (mesh->index_base != 0 && mesh->index_base != 1)) {
      set_error("lb_schedule_host: invalid argument");
      return LB_E_ARG;
    }
    lb_op op;
    op.host_only = true;
    op.NE = mesh->NE;
    op.NK = mesh->NK;
    op.NI = quad->NI;
    op.NENK = 4 * mesh->NE;
    op.L = quad->L;
    op.h_k.assign(quad->k, quad->k + 3 * quad->NI);
    op.h_w.assign(quad->w, quad->w + quad->NI);
    LB_TRY(compute_geometry(&op, mesh->rg, mesh->node, mesh->nbr, mesh->mat, mesh->index_base));
    LB_TRY(lb_op_get_schedule(&op, sl, nlev));
    if (normals) std::memcpy(normals, op.h_nrm.data(), op.h_nrm.size() * sizeof(double));
    return LB_OK;
  });
}

int lb_trace_plan_host(const lb_mesh_desc* mesh, const lb_quad_desc* quad, int G, int lds_bytes, int32_t* uslot,
                       int32_t* wslot, int32_t* stage_off, int32_t* stage, int64_t* nstage, int32_t* sizes,
                       int64_t* stats, uint8_t* gprod) {
  return guarded([&]() -> int {
    if (!mesh || !quad || !mesh->rg || !mesh->node || !mesh->nbr || !mesh->mat || !quad->k || !quad->w ||
        mesh->NE <= 0 || mesh->NE > LB_MAX_ELEMENTS || quad->NI <= 0 || G < 1 || lds_bytes <= 0 ||
        (mesh->index_base != 0 && mesh->index_base != 1)) {
      set_error("lb_trace_plan_host: invalid argument");
      return LB_E_ARG;
    }
    lb_op op;
    op.host_only = true;
    op.NE = mesh->NE;
    op.NK = mesh->NK;
    op.NI = quad->NI;
    op.NENK = 4 * mesh->NE;
    op.L = quad->L;
    op.h_k.assign(quad->k, quad->k + 3 * quad->NI);
    op.h_w.assign(quad->w, quad->w + quad->NI);
    LB_TRY(compute_geometry(&op, mesh->rg, mesh->node, mesh->nbr, mesh->mat, mesh->index_base));
    int ring = 0, st = 0;
    LB_TRY(trace_plan_sizes(lds_bytes, G, &ring, &st));
    const int64_t NI = op.NI, NE = op.NE, stride = op.maxlev + 2;
    std::vector<int32_t> us(NI * NE * 4), ws(NI * NE), so(NI * stride), ent;
    int64_t sts[4];
    LB_TRY(build_trace_plan(NI, NE, op.maxlev, op.h_nlev.data(), op.h_lvl_off.data(), op.h_tasks.data(),
                            op.h_nbr.data(), ring, st, us.data(), ws.data(), so.data(), ent, sts, gprod));
    if (uslot) std::memcpy(uslot, us.data(), us.size() * sizeof(int32_t));
    if (wslot) std::memcpy(wslot, ws.data(), ws.size() * sizeof(int32_t));
    if (stage_off) std::memcpy(stage_off, so.data(), so.size() * sizeof(int32_t));
    const int64_t n = (int64_t)ent.size() / 2;
    if (nstage) {
      if (stage && *nstage >= n) std::memcpy(stage, ent.data(), ent.size() * sizeof(int32_t));
      *nstage = n;
    }
    if (sizes) {
      sizes[0] = ring;
      sizes[1] = st;
    }
    if (stats) std::memcpy(stats, sts, sizeof(sts));
    return LB_OK;
  });
}

int lb_op_set_physics(lb_op* op, int Nm, const double* sigma_t, int L, const double* sigma_s) {
  return guarded([&]() -> int {
    return lb_op_set_physics_groups(op, 1, Nm, sigma_t, L, sigma_s);
  });
}

int lb_op_set_physics_groups(lb_op* op, int G, int Nm, const double* sigma_t, int L, const double* sigma_s) {
  return guarded([&]() -> int {
    if (!op || G < 1 || Nm < op->Nm || L < 0 || !sigma_t || !sigma_s) {
      set_error("lb_op_set_physics: need G >= 1, one sigma_t per mesh material (Nm >= " + std::to_string(op ? op->Nm : 0) +
                ") and L >= 0");
      return LB_E_ARG;
    }
    op->G = G;
    op->h_sigt.assign((size_t)op->Nm * G, 0.0);
    op->h_sigs.assign((size_t)op->Nm * G * (op->L + 1), 0.0);
    for (int g = 0; g < G; ++g)
      for (int m = 0; m < op->Nm; ++m) {
        op->h_sigt[(size_t)m * G + g] = sigma_t[(size_t)g * Nm + m];
        for (int l = 0; l <= std::min(L, op->L); ++l)
          op->h_sigs[((size_t)m * G + g) * (op->L + 1) + l] = sigma_s[((size_t)g * Nm + m) * (L + 1) + l];
      }
    return upload_physics(op);
  });
}

int lb_op_set_scatter_mode(lb_op* op, int mode) {
  return guarded([&]() -> int {
    if (!op || (mode != LB_SCATTER_MOMENT && mode != LB_SCATTER_DENSE)) {
      set_error("lb_op_set_scatter_mode: invalid mode");
      return LB_E_ARG;
    }
    op->scatter_mode = mode;
    return LB_OK;
  });
}

int lb_op_set_sweep_mode(lb_op* op, int mode) {
  return guarded([&]() -> int {
    if (!op || mode < LB_SWEEP_LEVELS || mode > LB_SWEEP_DIRECTION_LDS) {
      set_error("lb_op_set_sweep_mode: invalid mode");
      return LB_E_ARG;
    }
    if (op->stream) (void)hipStreamSynchronize(op->stream);
    clear_graphs(op);
    op->sweep_mode = mode;
    return LB_OK;
  });
}

int lb_op_sweep_choice(lb_op* op, int G, int* mode, int* threads, int* per_cu) {
  return guarded([&]() -> int {
    if (!op || !mode || !threads || !per_cu || G < 1) {
      set_error("lb_op_sweep_choice: invalid argument");
      return LB_E_ARG;
    }
    const bool batched = op->batch.P > 0;
    auto& tune = batched ? op->btune : op->tune;
    const SweepConfig* table = batched ? kBatchTuneConfig : kTuneConfig;
    auto it = tune.find(G);
    if (it == tune.end() || it->second.choice < 0) {
      *mode = LB_SWEEP_AUTO;
      *threads = 0;
      *per_cu = 0;
    } else {
      *mode = table[it->second.choice].mode;
      *threads = table[it->second.choice].block;
      *per_cu = table[it->second.choice].per_cu;
    }
    return LB_OK;
  });
}

int lb_op_set_group_batch(lb_op* op, int groups) {
  return guarded([&]() -> int {
    if (!op || groups < 1 || groups > kMaxGroupBatch) {
      set_error("lb_op_set_group_batch: groups must be 1..16");
      return LB_E_ARG;
    }
    op->group_batch = groups;
    return LB_OK;
  });
}

int lb_op_set_workgroups_per_cu(lb_op* op, int per_cu) {
  return guarded([&]() -> int {
    if (!op || per_cu < 1 || per_cu > 8) {
      set_error("lb_op_set_workgroups_per_cu: per_cu must be 1..8");
      return LB_E_ARG;
    }
    op->per_cu = per_cu;
    return LB_OK;
  });
}

int lb_op_set_problems(lb_op* op, int P, const int64_t* elem_off) {
  return guarded([&]() -> int {
    if (!op || op->host_only) {
      set_error("lb_op_set_problems: null operator");
      return LB_E_ARG;
    }
    return set_problems(op, P, elem_off);
  });
}

int lb_op_set_preconditioner(lb_op* op, const float* params, int nparams, int fmt) {
  return guarded([&]() -> int {
    if (!op || op->host_only) {
      set_error("lb_op_set_preconditioner: null operator");
      return LB_E_ARG;
    }
    return mlp_set(op, params, nparams, fmt);
  });
}

int lb_mlp_apply_dev(lb_op* op, const double* d_r, double* d_z, int g0) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_mlp_apply_dev"));
    if (!d_r || !d_z || g0 < 0 || g0 >= op->G) {
      set_error("lb_mlp_apply_dev: invalid argument");
      return LB_E_ARG;
    }
    return mlp_apply(op, d_r, d_z, g0);
  });
}

int lb_beam_uncollided_dev(lb_op* op, const lb_beam_desc* beam, const double* attn, double* d_phi0) {
  return guarded([&]() -> int {
    if (!op) {
      set_error("lb_beam_uncollided_dev: null operator");
      return LB_E_ARG;
    }
    return beam_uncollided(op, beam, attn, 0, op->NE, d_phi0);
  });
}

int lb_beam_uncollided_range_dev(lb_op* op, const lb_beam_desc* beam, const double* attn, int64_t e0, int64_t e1,
                                 double* d_phi0) {
  return guarded([&]() -> int {
    if (!op) {
      set_error("lb_beam_uncollided_range_dev: null operator");
      return LB_E_ARG;
    }
    return beam_uncollided(op, beam, attn, e0, e1, d_phi0);
  });
}

int lb_beam_moments_dev(lb_op* op, const lb_beam_desc* beam, const double* d_phi0, double* d_fcs, double* d_unc) {
  return guarded([&]() -> int {
    if (!op) {
      set_error("lb_beam_moments_dev: null operator");
      return LB_E_ARG;
    }
    return beam_moments(op, beam, d_phi0, d_fcs, d_unc);
  });
}

int lb_beam_angular_fluence_dev(lb_op* op, const lb_beam_desc* beam, const double* d_phi0, double* d_unc) {
  return guarded([&]() -> int {
    if (!op) {
      set_error("lb_beam_angular_fluence_dev: null operator");
      return LB_E_ARG;
    }
    return beam_angular_fluence(op, beam, d_phi0, d_unc);
  });
}

int lb_beam_source(lb_op* op, const lb_beam_desc* beam, double* s, int64_t n) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "ExternalBeam%Source"));
    LB_TRY(check_n(op, n, "ExternalBeam%Source"));
    if (!s) {
      set_error("ExternalBeam%Source: null source vector");
      return LB_E_ARG;
    }
    LB_TRY(to_device(op, s, n));
    LB_TRY(beam_source(op, beam, 0, op->w_a.as<double>()));
    return to_host(op, s, n);
  });
}

int lb_beam_add_angular_fluence(lb_op* op, const lb_beam_desc* beam, const double* attn, double* unc, int64_t n) {
  return guarded([&]() -> int {
    if (!op) {
      set_error("ExternalBeam%AddAngularFluence: null operator");
      return LB_E_ARG;
    }
    LB_TRY(check_n(op, n, "ExternalBeam%AddAngularFluence"));
    if (!unc || !attn) {
      set_error("ExternalBeam%AddAngularFluence: null argument");
      return LB_E_ARG;
    }
    LB_TRY(to_device(op, unc, n));
    LB_TRY(beam_add_angular_fluence(op, beam, attn, op->w_a.as<double>()));
    return to_host(op, unc, n);
  });
}

int lb_op_set_preconditioner_threshold(lb_op* op, double min_thickness) {
  return guarded([&]() -> int {
    if (!op || std::isnan(min_thickness) || min_thickness < 0.0) {
      set_error("lb_op_set_preconditioner_threshold: needs an operator and a non-negative thickness");
      return LB_E_ARG;
    }
    op->mlp_min_tau = min_thickness;
    return LB_OK;
  });
}

int lb_op_set_moment_precision(lb_op* op, int fmt) {
  return guarded([&]() -> int {
    if (!op || fmt < LB_FMT_FP32 || fmt > LB_FMT_FP64) {
      set_error("lb_op_set_moment_precision: needs an operator and an LB_FMT_* format");
      return LB_E_ARG;
    }
    if (fmt != LB_FMT_FP64 && (op->NI != 256 || op->nmom != 16)) {
      set_error("lb_op_set_moment_precision: reduced-precision products need 256 directions and P3");
      return LB_E_ARG;
    }
    op->moment_fmt = fmt;
    return LB_OK;
  });
}

int lb_mlp_apply_noalloc_dev(lb_op* op, const double* d_r, double* d_z, int g0) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_mlp_apply_noalloc_dev"));
    if (!d_r || !d_z || g0 < 0 || g0 >= op->G) {
      set_error("lb_mlp_apply_noalloc_dev: invalid argument");
      return LB_E_ARG;
    }
    return mlp_apply_noalloc(op, d_r, d_z, g0);
  });
}

int lb_mlp_features_dev(lb_op* op, const double* d_r, int g0, float* d_feat, double* scale) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_mlp_features_dev"));
    if (!d_r || !d_feat || g0 < 0 || g0 >= op->G) {
      set_error("lb_mlp_features_dev: invalid argument");
      return LB_E_ARG;
    }
    return mlp_features(op, d_r, g0, d_feat, scale);
  });
}

int lb_moment_rhs_dev(lb_op* op, const double* d_q, double* d_b, int g0) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_moment_rhs_dev"));
    if (!d_q || !d_b || g0 < 0 || g0 >= op->G) {
      set_error("lb_moment_rhs_dev: invalid argument");
      return LB_E_ARG;
    }
    return moment_rhs_dev(op, d_q, d_b, g0);
  });
}

int lb_discrete_source_dev(lb_op* op, const double* d_qmom, double* d_q) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_discrete_source_dev"));
    if (!d_qmom || !d_q) {
      set_error("lb_discrete_source_dev: null vector");
      return LB_E_ARG;
    }
    return launch_moments_to_discrete(op, d_qmom, d_q, 0, 1, 0, false);
  });
}

int lb_moment_matvec_dev(lb_op* op, const double* d_x, double* d_y, int g0) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_moment_matvec_dev"));
    if (!d_x || !d_y || g0 < 0 || g0 >= op->G) {
      set_error("lb_moment_matvec_dev: invalid argument");
      return LB_E_ARG;
    }
    if (d_x != d_y) LB_TRY(launch_copy(op->stream, op->nmom * op->NENK, d_x, d_y));
    return moment_matvec_dev(op, d_y, g0);
  });
}

int lb_gmres_snapshots_dev(lb_op* op, const double* d_q, const lb_solver_opts* opts, double* d_phi,
                           double* d_snaps, int max_snaps, int* nsnaps, lb_solve_stats* stats) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_gmres_snapshots_dev"));
    if (!d_q || !d_phi || (max_snaps > 0 && !d_snaps) || max_snaps < 0 || op->G != 1) {
      set_error("lb_gmres_snapshots_dev: invalid argument (single-group operators only)");
      return LB_E_ARG;
    }
    return gmres_snapshots_dev(op, d_q, opts, d_phi, d_snaps, max_snaps, nsnaps, stats);
  });
}

int lb_batch_solve_dev(lb_op* op, const double* d_qmom, const lb_solver_opts* opts, double* d_phi,
                       lb_solve_stats* stats) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_batch_solve_dev"));
    if (op->batch.P == 0) {
      set_error("lb_batch_solve_dev: no problem set (lb_op_set_problems)");
      return LB_E_ARG;
    }
    if (op->G != 1) {
      set_error("lb_batch_solve_dev: the batched solve is single-group");
      return LB_E_ARG;
    }
    if (!d_qmom || !d_phi) {
      set_error("lb_batch_solve_dev: null vector");
      return LB_E_ARG;
    }
    return batch_solve_dev(op, d_qmom, opts, d_phi, stats);
  });
}

int lb_op_set_direction_block(lb_op* op, int threads) {
  return guarded([&]() -> int {
    if (!op || (threads != 256 && threads != 512 && threads != 768 && threads != 1024)) {
      set_error("lb_op_set_direction_block: threads must be 256, 512, 768 or 1024");
      return LB_E_ARG;
    }
    op->dir_block = threads;
    return LB_OK;
  });
}

int lb_transport_matinv(lb_op* op, double* s, int64_t n) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "TransportOp%MatInv"));
    LB_TRY(check_n(op, n, "TransportOp%MatInv"));
    LB_TRY(to_device(op, s, n));
    LB_TRY(launch_sweep(op, op->w_a.as<double>(), op->w_a.as<double>(), 0, 1));
    return to_host(op, s, n);
  });
}

int lb_op_build_delta_down(lb_op* op, int Nm, const double* xsdd) {
  return guarded([&]() -> int {
    if (!op || !xsdd || Nm < op->Nm) {
      set_error("ScatteringOp%BuildDeltaDown: need one cross section per mesh material (Nm >= " +
                std::to_string(op ? op->Nm : 0) + ")");
      return LB_E_ARG;
    }
    LB_TRY(op->d_xsdd.reserve((size_t)op->Nm * sizeof(double)));
    LB_HIP(hipMemcpy(op->d_xsdd.p, xsdd, (size_t)op->Nm * sizeof(double), hipMemcpyHostToDevice));
    op->delta_down = true;
    return LB_OK;
  });
}

int lb_scatter_matvec(lb_op* op, double* x, int64_t n) {
  return guarded([&]() -> int {
    if (op && op->delta_down) {   // ApplyDeltaDown (ScatteringAlgebra.f90:62-90)
      LB_TRY(check_n(op, n, "ScatteringOp%MatVec"));
      LB_TRY(to_device(op, x, n));
      LB_TRY(launch_delta_down(op, n, op->w_a.as<double>()));
      return to_host(op, x, n);
    }
    LB_TRY(check_ready(op, "ScatteringOp%MatVec"));
    LB_TRY(check_n(op, n, "ScatteringOp%MatVec"));
    LB_TRY(to_device(op, x, n));
    LB_TRY(op->w_c.reserve(n * sizeof(double)));
    LB_TRY(scatter_dev(op, op->w_a.as<double>(), op->w_c.as<double>(), 0, 1));
    LB_HIP(hipMemcpyAsync(x, op->w_c.p, n * sizeof(double), hipMemcpyDeviceToHost, op->stream));
    LB_HIP(hipStreamSynchronize(op->stream));
    return LB_OK;
  });
}

int lb_boltzmann_matvec(lb_op* op, double* x, int64_t n) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "BoltzmannOp%MatVec"));
    LB_TRY(check_n(op, n, "BoltzmannOp%MatVec"));
    LB_TRY(to_device(op, x, n));
    LB_TRY(boltzmann_matvec_dev(op, op->w_a.as<double>(), 0));
    return to_host(op, x, n);
  });
}

int lb_boltzmann_matinv(lb_op* op, double* s, int64_t n, const lb_solver_opts* opts, lb_solve_stats* stats) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "BoltzmannOp%MatInv"));
    LB_TRY(check_n(op, n, "BoltzmannOp%MatInv"));
    LB_TRY(to_device(op, s, n));
    int rc = boltzmann_matinv_dev(op, op->w_a.as<double>(), 0, opts, stats);
    if (rc != LB_OK && rc != LB_E_MAXIT && rc != LB_E_BREAKDOWN && rc != LB_E_NAN) return rc;
    std::string err = g_error;
    LB_TRY(to_host(op, s, n));
    g_error = err;
    return rc;
  });
}

int lb_fluence(lb_op* op, const double* v, double* phi) {
  return guarded([&]() -> int {
    if (!op || !v || !phi) {
      set_error("lb_fluence: null argument");
      return LB_E_ARG;
    }
    const int64_t n = op->NENK * op->NI;
    LB_TRY(to_device(op, v, n));
    LB_TRY(op->w_c.reserve(op->NENK * sizeof(double)));
    LB_TRY(launch_fluence(op, op->w_a.as<double>(), op->w_c.as<double>()));
    LB_HIP(hipMemcpyAsync(phi, op->w_c.p, op->NENK * sizeof(double), hipMemcpyDeviceToHost, op->stream));
    LB_HIP(hipStreamSynchronize(op->stream));
    return LB_OK;
  });
}

int lb_op_get_schedule(lb_op* op, int32_t* sl, int32_t* nlev) {
  return guarded([&]() -> int {
    if (!op) {
      set_error("lb_op_get_schedule: null operator");
      return LB_E_ARG;
    }
    if (sl)
      for (size_t j = 0; j < op->h_tasks.size(); ++j) sl[j] = int32_t(op->h_tasks[j] & 0xFFFFFFu);
    if (nlev) std::memcpy(nlev, op->h_nlev.data(), op->h_nlev.size() * sizeof(int32_t));
    return LB_OK;
  });
}

int lb_op_get_normals(lb_op* op, double* n) {
  return guarded([&]() -> int {
    if (!op || !n) {
      set_error("lb_op_get_normals: null argument");
      return LB_E_ARG;
    }
    std::memcpy(n, op->h_nrm.data(), op->h_nrm.size() * sizeof(double));
    return LB_OK;
  });
}

int lb_op_info(lb_op* op, int64_t* NE, int64_t* NI, int64_t* nmom, int64_t* maxlev) {
  return guarded([&]() -> int {
    if (!op) {
      set_error("lb_op_info: null operator");
      return LB_E_ARG;
    }
    if (NE) *NE = op->NE;
    if (NI) *NI = op->NI;
    if (nmom) *nmom = op->nmom;
    if (maxlev) *maxlev = op->maxlev;
    return LB_OK;
  });
}

int lb_transport_matinv_dev(lb_op* op, const double* d_q, double* d_psi, int g0, int G) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_transport_matinv_dev"));
    if (g0 < 0 || G < 1 || g0 + G > op->G) {
      set_error("lb_transport_matinv_dev: group range outside the physics table");
      return LB_E_ARG;
    }
    return launch_sweep(op, d_q, d_psi, g0, G);
  });
}

int lb_moments_dev(lb_op* op, const double* d_psi, double* d_phi, int G) {
  return guarded([&]() -> int {
    if (!op || G < 1) {
      set_error("lb_moments_dev: invalid argument");
      return LB_E_ARG;
    }
    return launch_moments(op, d_psi, d_phi, G, op->NENK * G);
  });
}

int lb_op_set_flux_order(lb_op* op, int order) {
  return guarded([&]() -> int {
    if (!op || order < LB_FLUX_NATURAL || order > LB_FLUX_SWEEP_ORDER_FP32) {
      set_error("lb_op_set_flux_order: order must be LB_FLUX_NATURAL, LB_FLUX_SWEEP_ORDER, "
                "LB_FLUX_SWEEP_ORDER_SOURCE or LB_FLUX_SWEEP_ORDER_FP32");
      return LB_E_ARG;
    }
    op->flux_order = order;
    return LB_OK;
  });
}

int lb_op_sweep_order_choice(lb_op* op, int source_in_order, int* mode, int* threads) {
  return guarded([&]() -> int {
    if (!op || !mode || !threads) {
      set_error("lb_op_sweep_order_choice: invalid argument");
      return LB_E_ARG;
    }
    const bool batched = op->batch.P > 0;
    auto& tune = batched ? op->btune_pos : op->tune_pos;
    const SweepConfig* table = batched ? kBatchTuneConfig : kTuneConfig;
    auto it = tune.find(source_in_order ? kOrderBoth : kOrderFlux);
    if (it == tune.end() || it->second.choice < 0) {
      *mode = LB_SWEEP_AUTO;
      *threads = 0;
    } else {
      *mode = table[it->second.choice].mode;
      *threads = table[it->second.choice].block;
    }
    return LB_OK;
  });
}

int lb_transport_matinv_order_dev(lb_op* op, const double* d_q, double* d_psi_so, int g0, int source_in_order) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_transport_matinv_order_dev"));
    if (g0 < 0 || g0 >= op->G || !d_q || !d_psi_so) {
      set_error("lb_transport_matinv_order_dev: invalid argument");
      return LB_E_ARG;
    }
    const int64_t n = op->NI * op->NENK;
    const bool same = source_in_order && d_q == d_psi_so;
    if (!same && d_q < d_psi_so + n && d_psi_so < d_q + n) {
      set_error("lb_transport_matinv_order_dev: the source and the sweep-order flux overlap");
      return LB_E_ARG;
    }
    return launch_sweep_pos(op, d_q, d_psi_so, g0, source_in_order != 0);
  });
}

int lb_moments_to_discrete_order_dev(lb_op* op, const double* d_phi, double* d_psi_so, int g0) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_moments_to_discrete_order_dev"));
    if (g0 < 0 || g0 >= op->G || !d_phi || !d_psi_so) {
      set_error("lb_moments_to_discrete_order_dev: invalid argument");
      return LB_E_ARG;
    }
    return launch_moments_to_order(op, d_phi, d_psi_so, g0);
  });
}

int lb_transport_matinv_order32_dev(lb_op* op, const double* d_q, double* d_psi_so, float* d_psi32_so, int g0) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_transport_matinv_order32_dev"));
    if (g0 < 0 || g0 >= op->G || !d_q || !d_psi_so || !d_psi32_so) {
      set_error("lb_transport_matinv_order32_dev: invalid argument");
      return LB_E_ARG;
    }
    const int64_t n = op->NI * op->NENK;
    const char* qa = reinterpret_cast<const char*>(d_q);
    const char* pa = reinterpret_cast<const char*>(d_psi_so);
    const char* fa = reinterpret_cast<const char*>(d_psi32_so);
    auto overlap = [](const char* a, size_t na, const char* b, size_t nb) { return a < b + nb && b < a + na; };
    if (overlap(qa, n * 8, pa, n * 8) || overlap(qa, n * 8, fa, n * 4) || overlap(pa, n * 8, fa, n * 4)) {
      set_error("lb_transport_matinv_order32_dev: the source and the outputs overlap");
      return LB_E_ARG;
    }
    return launch_sweep_f32(op, d_q, d_psi_so, d_psi32_so, g0);
  });
}

int lb_moments_order32_dev(lb_op* op, const float* d_psi32_so, double* d_phi) {
  return guarded([&]() -> int {
    if (!op || !d_psi32_so || !d_phi) {
      set_error("lb_moments_order32_dev: invalid argument");
      return LB_E_ARG;
    }
    return launch_moments_pos32(op, d_psi32_so, d_phi);
  });
}

int lb_moments_order_dev(lb_op* op, const double* d_psi_so, double* d_phi) {
  return guarded([&]() -> int {
    if (!op || !d_psi_so || !d_phi) {
      set_error("lb_moments_order_dev: invalid argument");
      return LB_E_ARG;
    }
    return launch_moments_pos(op, d_psi_so, d_phi);
  });
}

int lb_flux_to_natural_dev(lb_op* op, const double* d_psi_so, double* d_psi) {
  return guarded([&]() -> int {
    if (!op || !d_psi_so || !d_psi) {
      set_error("lb_flux_to_natural_dev: invalid argument");
      return LB_E_ARG;
    }
    const int64_t n = op->NI * op->NENK;
    if (d_psi_so < d_psi + n && d_psi < d_psi_so + n) {
      set_error("lb_flux_to_natural_dev: the two layouts' vectors overlap");
      return LB_E_ARG;
    }
    return launch_flux_to_natural(op, d_psi_so, d_psi);
  });
}

int lb_moments_to_discrete_dev(lb_op* op, const double* d_phi, double* d_psi, int g0, int G, int accumulate) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_moments_to_discrete_dev"));
    if (g0 < 0 || G < 1 || g0 + G > op->G) {
      set_error("lb_moments_to_discrete_dev: group range outside the physics table");
      return LB_E_ARG;
    }
    return launch_moments_to_discrete(op, d_phi, d_psi, g0, G, accumulate);
  });
}

int lb_multigroup_solve_dev(lb_op* op, const double* d_qmom, const double* transfer, const lb_solver_opts* opts,
                            double* d_phi, lb_solve_stats* stats) {
  return guarded([&]() -> int {
    LB_TRY(check_ready(op, "lb_multigroup_solve_dev"));
    if (!transfer || !d_phi) {
      set_error("lb_multigroup_solve_dev: null argument");
      return LB_E_ARG;
    }
    return multigroup_solve_dev(op, d_qmom, transfer, opts, d_phi, stats);
  });
}

int lb_op_synchronize(lb_op* op) {
  return guarded([&]() -> int {
    if (!op) {
      set_error("lb_op_synchronize: null operator");
      return LB_E_ARG;
    }
    LB_HIP(hipStreamSynchronize(op->stream));
    return LB_OK;
  });
}

int lb_last_sweep_ms(lb_op* op, double* ms) {
  return guarded([&]() -> int {
    if (!op || !ms) {
      set_error("lb_last_sweep_ms: null argument");
      return LB_E_ARG;
    }
    float f = 0.f;
    LB_HIP(hipEventSynchronize(op->ev1));
    LB_HIP(hipEventElapsedTime(&f, op->ev0, op->ev1));
    *ms = f;
    return LB_OK;
  });
}

void lb_op_destroy(lb_op* op) {
  if (!op) return;
  if (op->stream) (void)hipStreamSynchronize(op->stream);
  clear_graphs(op);
  for (DevBuf* b : {&op->d_geo, &op->d_nbr, &op->d_perm, &op->d_mat, &op->d_tasks, &op->d_k, &op->d_w, &op->d_D,
                    &op->d_M, &op->d_deg, &op->d_yys, &op->d_sigt, &op->d_sigs, &op->w_a, &op->w_a2, &op->w_b,
                    &op->w_c, &op->w_phi, &op->w_red, &op->w_krylov, &op->w_mg, &op->w_acc, &op->w_S,
                    &op->w_arn})
    b->release();
  for (auto& kv : op->tables) kv.second.d_entries.release();
  if (op->w_host.p) (void)hipHostFree(op->w_host.p);
  op->d_rec.release();
  op->d_rec_pos.release();
  op->d_inv.release();
  op->w_so.release();
  op->w_so32.release();
  for (DevBuf* b : {&op->batch.d_items, &op->batch.d_lvl, &op->batch.d_tasks, &op->batch.d_rec,
                    &op->batch.d_inv, &op->batch.d_rec_pos,
                    &op->batch.d_active, &op->batch.d_counter, &op->batch.d_coef, &op->batch.d_part,
                    &op->batch.d_out})
    b->release();
  for (DevBuf* b : {&op->d_mlp_w, &op->d_mlp_pack, &op->d_elem_h, &op->w_feat, &op->d_vert, &op->w_beam, &op->w_mix,
                    &op->w_phi0})
    b->release();
  for (auto& kv : op->batch.seg) {
    kv.second.d_tiles.release();
    kv.second.d_tile_off.release();
  }
  if (op->batch.h_io) (void)hipHostFree(op->batch.h_io);
  op->batch.h_io = nullptr;
  op->d_lvl_off.release();
  op->d_nlev.release();
  for (auto& kv : op->plans)
    for (DevBuf* b : {&kv.second.d_uslot, &kv.second.d_wslot, &kv.second.d_stage_off, &kv.second.d_stage,
                      &kv.second.d_stage_pos, &kv.second.d_gprod})
      b->release();
  op->d_xsdd.release();
  op->w_phis.release();
  op->w_tq.release();
  op->w_tail.release();
  release_blas(op);
  if (op->ev0) (void)hipEventDestroy(op->ev0);
  if (op->ev1) (void)hipEventDestroy(op->ev1);
  if (op->stream) (void)hipStreamDestroy(op->stream);
  delete op;
}

}  // extern "C"
