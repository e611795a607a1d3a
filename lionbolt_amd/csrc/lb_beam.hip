// Planar external beam: uncollided fluence and first-collision source moments.
//
// The reference (sources/) builds the first-collision source of a beam from the
// uncollided fluence at the four tetrahedral quadrature points of every element
// in the field: phi0q = n exp(-sum_m sigma_t,m l_m), l_m the path length in
// material m from the point back along the beam axis to the mesh surface
// (RayTracing.f90:149-199, FluenceQuadrature.f90:3-33), projected on the nodal
// basis, phi0 = IMr (sum_kq w_kq u_k(kq) phi0q) (FCSSource.f90:134-176; the same
// in AddAngularFluence.f90:109-151).  The reference intersects each ray with
// every face enclosing each material region; here one thread per quadrature
// point walks the ray through the tetrahedra instead, element to element across
// the exit face (the smallest positive distance to a face plane along the ray),
// adding sigma_t of each element times the length inside it.  A straight ray
// leaves each convex element once, so the walk ends within NE steps.  The four
// threads of an element then exchange their values by lane shuffles for the
// projection.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "lb_internal.hpp"

namespace lb {

namespace {

// The tetrahedral quadrature: barycentric weights (a, b, b, b) and their
// permutations, point kq nearest node kq, weight 1/24 on the reference element.
constexpr double kQa = 0.5854101966249685;   // (5 + 3 sqrt 5) / 20
constexpr double kQb = 0.1381966011250105;   // (5 - sqrt 5) / 20

struct BeamArgs {
  const double* vert;    // [NE][4][3]
  const double4* geo;    // [NE][4]: outward unit normal, c_f
  const int4* nbr;       // [NE], -1 on the boundary
  const int32_t* mat;    // [NE]
  const double* attn;    // [Nm] sigma_t per material
  double axis[3], origin[3], e1[3], e2[3];
  int cutout;
  double params[4];
  double n;
  int64_t NE;
  int64_t elem0, elem1;  // elements traced: [elem0, elem1)
  int max_steps;
};

// InField.f90:3-62, planar beams (cutout none, circle, rectangle in the beam frame).
__device__ bool beam_in_field(const BeamArgs& B, double x, double y, double z) {
  if (B.cutout == LB_CUTOUT_NONE) return true;
  const double ux = x - B.origin[0], uy = y - B.origin[1], uz = z - B.origin[2];
  const double ua = ux * B.axis[0] + uy * B.axis[1] + uz * B.axis[2];
  const double vx = ux - ua * B.axis[0], vy = uy - ua * B.axis[1], vz = uz - ua * B.axis[2];
  constexpr double kTol = 1.0e-12;
  if (B.cutout == LB_CUTOUT_CIRCLE) return sqrt(vx * vx + vy * vy + vz * vz) <= B.params[1] + kTol;
  if (!(fabs(vx * B.e1[0] + vy * B.e1[1] + vz * B.e1[2]) <= 0.5 * B.params[2] + kTol)) return false;
  return fabs(vx * B.e2[0] + vy * B.e2[1] + vz * B.e2[2]) <= 0.5 * B.params[3] + kTol;
}

__global__ __launch_bounds__(256) void beam_uncollided_kernel(BeamArgs B, double* __restrict__ phi0,
                                                              int* __restrict__ fails) {
  const int64_t q = 4 * B.elem0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool valid = q < 4 * B.elem1;
  const int64_t e = valid ? q >> 2 : 0;
  const int kq = int(q & 3);
  double phi = 0.0;
  if (valid) {
    const double* V = B.vert + e * 12;
    double px = 0.0, py = 0.0, pz = 0.0;   // MapToTetrahedralQuad
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double w = k == kq ? kQa : kQb;
      px += w * V[3 * k];
      py += w * V[3 * k + 1];
      pz += w * V[3 * k + 2];
    }
    if (beam_in_field(B, px, py, pz)) {
      const double dx = -B.axis[0], dy = -B.axis[1], dz = -B.axis[2];
      int64_t cur = e;
      double tau = 0.0;
      bool ok = false;
      for (int step = 0; step <= B.max_steps; ++step) {
        double bt = 1.0e300;
        int bf = -1;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          const double4 g = B.geo[cur * 4 + f];
          const double dn = g.x * dx + g.y * dy + g.z * dz;
          if (dn > 0.0) {
            const double* P0 = B.vert + cur * 12 + 3 * (f < 3 ? 0 : 1);   // a node of local face f
            const double t = (g.x * (P0[0] - px) + g.y * (P0[1] - py) + g.z * (P0[2] - pz)) / dn;
            if (t < bt) {
              bt = t;
              bf = f;
            }
          }
        }
        if (bf < 0) break;   // no face ahead: a degenerate element
        const double t = bt > 0.0 ? bt : 0.0;
        tau += B.attn[B.mat[cur]] * t;
        px += t * dx;
        py += t * dy;
        pz += t * dz;
        const int4 nb = B.nbr[cur];
        const int nx = bf == 0 ? nb.x : (bf == 1 ? nb.y : (bf == 2 ? nb.z : nb.w));
        if (nx < 0) {
          ok = true;
          break;
        }
        cur = nx;
      }
      if (ok)
        phi = B.n * exp(-tau);
      else
        atomicAdd(fails, 1);
    }
  }
  // projection: acc_k = sum_kq (1/24) u_k(kq) phi_kq, then IMr = 120 (I - J/5)
  const int base = (int)(threadIdx.x & 63) & ~3;
  double p[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) p[j] = __shfl(phi, base + j, 64);
  double acc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    acc[k] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[k] += (1.0 / 24.0) * (k == j ? kQa : kQb) * p[j];
  }
  if (valid) phi0[q] = 120.0 * (acc[kq] - (acc[0] + acc[1] + acc[2] + acc[3]) / 5.0);
}

// out[r][sd] = coef[r] phi0[sd] for the moments r < nmom.
__global__ __launch_bounds__(256) void beam_outer_kernel(int nmom, int64_t NENK, const double* __restrict__ coef,
                                                         const double* __restrict__ phi0, double* __restrict__ out) {
  const int64_t n = (int64_t)nmom * NENK;
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    const int r = int(j / NENK);
    out[j] = coef[r] * phi0[j - (int64_t)r * NENK];
  }
}

// unc[i][sd] += g[i] phi0[sd].
__global__ __launch_bounds__(256) void beam_angular_kernel(int64_t NI, int64_t NENK, const double* __restrict__ g,
                                                           const double* __restrict__ phi0, double* __restrict__ unc) {
  const int64_t n = NI * NENK;
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    const int64_t i = j / NENK;
    unc[j] += g[i] * phi0[j - i * NENK];
  }
}

// s[i][sd] += fac[i][mat(sd)] phi0[sd]: the first-collision source of
// FCSSource.f90:178-215, fac the angular factor of each (direction, material).
__global__ __launch_bounds__(256) void beam_source_kernel(int64_t NI, int64_t NENK, int Nm,
                                                          const double* __restrict__ fac,
                                                          const int32_t* __restrict__ mat,
                                                          const double* __restrict__ phi0, double* __restrict__ s) {
  const int64_t n = NI * NENK;
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    const int64_t i = j / NENK, sd = j - i * NENK;
    s[j] += fac[i * Nm + mat[sd >> 2]] * phi0[sd];
  }
}

__global__ void zero_int_kernel(int* x) {
  if (threadIdx.x == 0) x[0] = 0;
}

unsigned stride_grid(int64_t n) {
  const int64_t g = (n + 255) / 256;
  return (unsigned)(g < 8192 ? (g > 0 ? g : 1) : 8192);
}

int check_beam(const lb_op* op, const lb_beam_desc* b) {
  if (!b) {
    set_error("beam: null description");
    return LB_E_ARG;
  }
  const double a2 = b->axis[0] * b->axis[0] + b->axis[1] * b->axis[1] + b->axis[2] * b->axis[2];
  if (!(std::fabs(a2 - 1.0) <= 1e-12)) {
    set_error("beam: the axis must be a unit vector");
    return LB_E_ARG;
  }
  if (b->cutout < LB_CUTOUT_NONE || b->cutout > LB_CUTOUT_RECTANGLE) {
    set_error("beam: unknown cutout");
    return LB_E_ARG;
  }
  (void)op;
  return LB_OK;
}

// g_i = sum_l (2l+1) P_l(k_i . axis) / (4 pi): GenAngularDistribution (Sources.f90:257-276).
std::vector<double> angular_distribution(const lb_op* op, const double* axis) {
  std::vector<double> g(op->NI);
  for (int64_t i = 0; i < op->NI; ++i) {
    const double* k = &op->h_k[i * 3];
    const double x = k[0] * axis[0] + k[1] * axis[1] + k[2] * axis[2];
    double p0 = 1.0, p1 = x, f = 0.0;
    for (int l = 0; l <= op->L; ++l) {
      double P;
      if (l == 0) P = 1.0;
      else if (l == 1) P = x;
      else {
        P = ((2 * l - 1) * x * p1 - (l - 1) * p0) / l;
        p0 = p1;
        p1 = P;
      }
      f += (2 * l + 1) * P / (4.0 * M_PI);
    }
    g[i] = f;
  }
  return g;
}

}  // namespace

int beam_uncollided(lb_op* op, const lb_beam_desc* beam, const double* attn, int64_t e0, int64_t e1,
                    double* d_phi0) {
  LB_TRY(check_beam(op, beam));
  if (e0 < 0 || e1 > op->NE || e0 >= e1) {
    set_error("lb_beam_uncollided_dev: element range outside the mesh or empty");
    return LB_E_ARG;
  }
  if (!attn || !d_phi0) {
    set_error("lb_beam_uncollided_dev: null argument");
    return LB_E_ARG;
  }
  for (int m = 0; m < op->Nm; ++m)
    if (!std::isfinite(attn[m]) || attn[m] < 0.0) {
      set_error("lb_beam_uncollided_dev: attenuation must be finite and non-negative");
      return LB_E_ARG;
    }
  if (!op->d_vert.p) LB_TRY(upload(op, op->d_vert, op->h_vert.data(), op->h_vert.size() * sizeof(double)));
  LB_TRY(op->w_beam.reserve((size_t)(op->Nm + 8) * sizeof(double)));
  double* d_attn = op->w_beam.as<double>();
  int* d_fail = reinterpret_cast<int*>(d_attn + op->Nm);
  LB_HIP(hipMemcpyAsync(d_attn, attn, op->Nm * sizeof(double), hipMemcpyHostToDevice, op->stream));
  BeamArgs B{};
  B.vert = op->d_vert.as<const double>();
  B.geo = op->d_geo.as<const double4>();
  B.nbr = op->d_nbr.as<const int4>();
  B.mat = op->d_mat.as<const int32_t>();
  B.attn = d_attn;
  for (int c = 0; c < 3; ++c) {
    B.axis[c] = beam->axis[c];
    B.origin[c] = beam->origin[c];
    B.e1[c] = beam->e1[c];
    B.e2[c] = beam->e2[c];
  }
  B.cutout = beam->cutout;
  for (int c = 0; c < 4; ++c) B.params[c] = beam->params[c];
  B.n = beam->n;
  B.NE = op->NE;
  B.elem0 = e0;
  B.elem1 = e1;
  B.max_steps = (int)std::min<int64_t>(op->NE, 1 << 30);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, op->stream, d_fail);
  hipLaunchKernelGGL(beam_uncollided_kernel, dim3((unsigned)((4 * (e1 - e0) + 255) / 256)), dim3(256), 0, op->stream, B,
                     d_phi0, d_fail);
  LB_HIP(hipGetLastError());
  int fails = 0;
  LB_HIP(hipMemcpyAsync(&fails, d_fail, sizeof(int), hipMemcpyDeviceToHost, op->stream));
  LB_HIP(hipStreamSynchronize(op->stream));
  if (fails) {
    set_error("lb_beam_uncollided_dev: " + std::to_string(fails) +
              " rays found no way out of the mesh (degenerate elements?)");
    return LB_E_ARG;
  }
  return LB_OK;
}

int beam_moments(lb_op* op, const lb_beam_desc* beam, const double* d_phi0, double* d_fcs, double* d_unc) {
  LB_TRY(check_beam(op, beam));
  if (!d_phi0) {
    set_error("lb_beam_moments_dev: null argument");
    return LB_E_ARG;
  }
  const int nm = op->nmom;
  std::vector<double> y(nm), c(nm, 0.0);
  real_sph_harm(op->L, beam->axis, y.data());
  const std::vector<double> g = angular_distribution(op, beam->axis);
  for (int r = 0; r < nm; ++r)   // (D g)_r = sum_i w_i Y_r(k_i) g_i
    for (int64_t i = 0; i < op->NI; ++i) c[r] += op->h_D[(size_t)r * op->NI + i] * g[i];
  LB_TRY(op->w_beam.reserve((size_t)(op->Nm + 8 + 2 * nm) * sizeof(double)));
  double* d_y = op->w_beam.as<double>() + op->Nm + 8;
  double* d_c = d_y + nm;
  LB_HIP(hipMemcpy(d_y, y.data(), nm * sizeof(double), hipMemcpyHostToDevice));
  LB_HIP(hipMemcpy(d_c, c.data(), nm * sizeof(double), hipMemcpyHostToDevice));
  const int64_t n = (int64_t)nm * op->NENK;
  if (d_fcs)
    hipLaunchKernelGGL(beam_outer_kernel, dim3(stride_grid(n)), dim3(256), 0, op->stream, nm, op->NENK, d_y, d_phi0,
                       d_fcs);
  if (d_unc)
    hipLaunchKernelGGL(beam_outer_kernel, dim3(stride_grid(n)), dim3(256), 0, op->stream, nm, op->NENK, d_c, d_phi0,
                       d_unc);
  LB_HIP(hipGetLastError());
  LB_HIP(hipStreamSynchronize(op->stream));
  return LB_OK;
}

int beam_angular_fluence(lb_op* op, const lb_beam_desc* beam, const double* d_phi0, double* d_unc) {
  LB_TRY(check_beam(op, beam));
  if (!d_phi0 || !d_unc) {
    set_error("lb_beam_angular_fluence_dev: null argument");
    return LB_E_ARG;
  }
  const std::vector<double> g = angular_distribution(op, beam->axis);
  LB_TRY(op->w_beam.reserve((size_t)(op->Nm + 8 + std::max<int64_t>(2 * op->nmom, op->NI)) * sizeof(double)));
  double* d_g = op->w_beam.as<double>() + op->Nm + 8;
  LB_HIP(hipMemcpy(d_g, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice));
  const int64_t n = op->NI * op->NENK;
  hipLaunchKernelGGL(beam_angular_kernel, dim3(stride_grid(n)), dim3(256), 0, op->stream, op->NI, op->NENK, d_g,
                     d_phi0, d_unc);
  LB_HIP(hipGetLastError());
  LB_HIP(hipStreamSynchronize(op->stream));
  return LB_OK;
}

int beam_source(lb_op* op, const lb_beam_desc* beam, int g0, double* d_s) {
  LB_TRY(check_beam(op, beam));
  if (g0 < 0 || g0 >= op->G) {
    set_error("lb_beam_source: group outside the operator's physics");
    return LB_E_ARG;
  }
  // attenuation sigma_t(mat, g0), as FluenceQuadrature(OpK%XS%t) (FCSSource.f90:47)
  std::vector<double> attn(op->Nm);
  for (int m = 0; m < op->Nm; ++m) attn[m] = op->h_sigt[(size_t)m * op->G + g0];
  LB_TRY(op->w_phi0.reserve((size_t)op->NENK * sizeof(double)));
  double* d_phi0 = op->w_phi0.as<double>();
  LB_TRY(beam_uncollided(op, beam, attn.data(), 0, op->NE, d_phi0));
  // fac(i, m) = sum_l (2l+1)/(4 pi) P_l(k_i . axis) sigma_l(m, g0) (FCSSource.f90:184-211)
  std::vector<double> fac((size_t)op->NI * op->Nm, 0.0);
  for (int64_t i = 0; i < op->NI; ++i) {
    const double* k = &op->h_k[i * 3];
    const double x = k[0] * beam->axis[0] + k[1] * beam->axis[1] + k[2] * beam->axis[2];
    double p0 = 1.0, p1 = x;
    for (int l = 0; l <= op->L; ++l) {
      double P;
      if (l == 0) P = 1.0;
      else if (l == 1) P = x;
      else {
        P = ((2 * l - 1) * x * p1 - (l - 1) * p0) / l;
        p0 = p1;
        p1 = P;
      }
      for (int m = 0; m < op->Nm; ++m)
        fac[(size_t)i * op->Nm + m] +=
            (2 * l + 1) * P / (4.0 * M_PI) * op->h_sigs[((size_t)m * op->G + g0) * (op->L + 1) + l];
    }
  }
  LB_TRY(op->w_beam.reserve((size_t)(op->Nm + 8 + fac.size()) * sizeof(double)));
  double* d_fac = op->w_beam.as<double>() + op->Nm + 8;
  LB_HIP(hipMemcpy(d_fac, fac.data(), fac.size() * sizeof(double), hipMemcpyHostToDevice));
  const int64_t n = op->NI * op->NENK;
  hipLaunchKernelGGL(beam_source_kernel, dim3(stride_grid(n)), dim3(256), 0, op->stream, op->NI, op->NENK, op->Nm,
                     d_fac, op->d_mat.as<const int32_t>(), d_phi0, d_s);
  LB_HIP(hipGetLastError());
  LB_HIP(hipStreamSynchronize(op->stream));
  return LB_OK;
}

int beam_add_angular_fluence(lb_op* op, const lb_beam_desc* beam, const double* attn, double* d_unc) {
  LB_TRY(op->w_phi0.reserve((size_t)op->NENK * sizeof(double)));
  double* d_phi0 = op->w_phi0.as<double>();
  LB_TRY(beam_uncollided(op, beam, attn, 0, op->NE, d_phi0));
  return beam_angular_fluence(op, beam, d_phi0, d_unc);
}

}  // namespace lb
