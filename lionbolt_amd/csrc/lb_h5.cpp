// liblbh5: results in the reference's HDF5 layout (include/lionbolt_h5.h) over
// the HDF5 C library.
#include <hdf5.h>

#include <algorithm>
#include <exception>
#include <cctype>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lionbolt_h5.h"

struct lbh5 {
  hid_t file = -1;
};

namespace {

thread_local std::string g_err;

// Entry points run their bodies through guarded(): a C++ exception (a failed
// host allocation, say) becomes a status code instead of crossing the C ABI.
template <class F>
int guarded(F&& f) noexcept {
  try {
    return f();
  } catch (...) {
    try {
      g_err = "unexpected C++ exception (host memory allocation failed?)";
    } catch (...) {
    }
    return -1;
  }
}

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

std::string lower(const char* s) {
  std::string r(s);
  std::transform(r.begin(), r.end(), r.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  return r;
}

// Opens the group at loc/name, creating it when absent.
hid_t open_group(hid_t loc, const std::string& name, bool* created = nullptr) {
  if (created) *created = false;
  htri_t ex = H5Lexists(loc, name.c_str(), H5P_DEFAULT);
  if (ex > 0) return H5Gopen2(loc, name.c_str(), H5P_DEFAULT);
  if (created) *created = true;
  return H5Gcreate2(loc, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
}

int attr_int(hid_t loc, const char* name, int32_t v) {
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(loc, name, H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT);
  herr_t e = a >= 0 ? H5Awrite(a, H5T_NATIVE_INT32, &v) : -1;
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  return e < 0 ? fail(std::string("attribute ") + name) : 0;
}

int attr_str(hid_t loc, const char* name, const std::string& v) {
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, std::max<size_t>(v.size(), 1));
  H5Tset_strpad(t, H5T_STR_NULLPAD);
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t a = H5Acreate2(loc, name, t, sp, H5P_DEFAULT, H5P_DEFAULT);
  herr_t e = a >= 0 ? H5Awrite(a, t, v.c_str()) : -1;
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  H5Tclose(t);
  return e < 0 ? fail(std::string("attribute ") + name) : 0;
}

template <class T>
int dset(hid_t loc, const std::string& name, hid_t file_type, hid_t mem_type, int rank, const hsize_t* dims,
         const T* data) {
  hid_t sp = H5Screate_simple(rank, dims, nullptr);
  hid_t d = H5Dcreate2(loc, name.c_str(), file_type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  herr_t e = d >= 0 ? H5Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) : -1;
  if (d >= 0) H5Dclose(d);
  H5Sclose(sp);
  return e < 0 ? fail("dataset " + name) : 0;
}

int dset_f64(hid_t loc, const std::string& name, int rank, const hsize_t* dims, const double* data) {
  return dset(loc, name, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, rank, dims, data);
}

int dset_i32(hid_t loc, const std::string& name, int64_t n, const int32_t* data) {
  const hsize_t d[1] = {(hsize_t)n};
  return dset(loc, name, H5T_STD_I32LE, H5T_NATIVE_INT32, 1, d, data);
}

bool fits_i32(int64_t v) { return v >= 0 && v <= INT32_MAX; }

}  // namespace

extern "C" {

const char* lbh5_last_error(void) { return g_err.c_str(); }

int lbh5_create(const char* path, lbh5** out) {
  return guarded([&]() -> int {
    if (!path || !out) return fail("lbh5_create: null argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors come back as return codes
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(std::string("lbh5_create: cannot create ") + path);
    *out = new lbh5{f};
    return 0;
  });
}

int lbh5_write_problem_type(lbh5* h, const char* problem_type) {
  return guarded([&]() -> int {
    if (!h || !problem_type) return fail("lbh5_write_problem_type: null argument");
    return attr_str(h->file, "problem_type", problem_type);
  });
}

int lbh5_write_mesh(lbh5* h, int64_t NE, int64_t NK, const int64_t* node, const double* rg, const int64_t* mat,
                    const double* vol, int slab) {
  return guarded([&]() -> int {
    if (!h || !node || !rg || !mat || !vol || NE <= 0 || NK <= 0) return fail("lbh5_write_mesh: invalid argument");
    const int64_t NENK = 4 * NE;
    if (!fits_i32(NENK) || !fits_i32(NK)) return fail("lbh5_write_mesh: sizes exceed 32-bit integers");
    int32_t Nm = 0;
    for (int64_t e = 0; e < NE; ++e) {
      if (mat[e] < 1) return fail("lbh5_write_mesh: materials are 1-based");
      Nm = std::max<int32_t>(Nm, (int32_t)mat[e]);
    }
    hid_t g = open_group(h->file, "mesh");
    if (g < 0) return fail("lbh5_write_mesh: group mesh");
    int rc = 0;
    rc = rc ? rc : attr_str(g, "problem_type", slab ? "slab" : "general");
    rc = rc ? rc : attr_int(g, "number_of_materials", Nm);
    rc = rc ? rc : attr_int(g, "number_of_elements", (int32_t)NE);
    rc = rc ? rc : attr_int(g, "number_of_spatial_dofs", (int32_t)NENK);
    rc = rc ? rc : attr_int(g, "number_of_global_nodes", (int32_t)NK);
    std::vector<int32_t> conn(NENK), off(NE + 1);
    for (int64_t e = 0; e < NE; ++e) {
      off[e] = (int32_t)(4 * e);
      for (int k = 0; k < 4; ++k) conn[4 * e + k] = (int32_t)(node[4 * e + k] + 1);
    }
    off[NE] = (int32_t)NENK;
    rc = rc ? rc : dset_i32(g, "connectivity", NENK, conn.data());
    rc = rc ? rc : dset_i32(g, "offset", NE + 1, off.data());
    const hsize_t rdims[2] = {(hsize_t)NK, 3};
    rc = rc ? rc : dset_f64(g, "global_nodes", 2, rdims, rg);
    if (!rc) {
      hid_t mg = open_group(g, "material_to_spatial_dof");
      if (mg < 0) {
        rc = fail("lbh5_write_mesh: group material_to_spatial_dof");
      } else {
        for (int32_t m = 1; m <= Nm && !rc; ++m) {
          std::vector<int32_t> sd;
          for (int64_t e = 0; e < NE; ++e)
            if (mat[e] == m)
              for (int k = 0; k < 4; ++k) sd.push_back((int32_t)(4 * e + k + 1));
          rc = dset_i32(mg, "material_" + std::to_string(m), (int64_t)sd.size(), sd.data());
        }
        H5Gclose(mg);
      }
    }
    const hsize_t vdims[1] = {(hsize_t)NE};
    rc = rc ? rc : dset_f64(g, "element_volumes", 1, vdims, vol);
    H5Gclose(g);
    return rc;
  });
}

int lbh5_write_particle(lbh5* h, const char* particle, int64_t NI, const double* w, const double* k, int G,
                        const double* E) {
  return guarded([&]() -> int {
    if (!h || !particle || !w || !k || NI <= 0 || !fits_i32(NI) || G < 0 || (G > 0 && !E))
      return fail("lbh5_write_particle: invalid argument");
    bool created = false;
    hid_t p = open_group(h->file, lower(particle), &created);
    if (p < 0) return fail("lbh5_write_particle: particle group");
    int rc = 0;
    if (created) rc = attr_str(p, "object_type", "particle");
    rc = rc ? rc : attr_int(p, "number_of_angular_dofs", (int32_t)NI);
    const hsize_t wd[1] = {(hsize_t)NI}, kd[2] = {(hsize_t)NI, 3};
    rc = rc ? rc : dset_f64(p, "angular_weights", 1, wd, w);
    rc = rc ? rc : dset_f64(p, "abscissae", 2, kd, k);
    if (G > 0) {
      const hsize_t ed[1] = {(hsize_t)(G + 1)};
      rc = rc ? rc : attr_int(p, "number_of_energy_dofs", G);
      rc = rc ? rc : dset_f64(p, "energy_grid", 1, ed, E);
    }
    H5Gclose(p);
    return rc;
  });
}

int lbh5_append_angular_fluence(lbh5* h, const char* particle, int uncollided, int g, int64_t NI, int64_t NENK,
                                const double* v) {
  return guarded([&]() -> int {
    if (!h || !particle || !v || g < 1 || NI <= 0 || NENK <= 0)
      return fail("lbh5_append_angular_fluence: invalid argument");
    hid_t p = open_group(h->file, lower(particle));
    if (p < 0) return fail("lbh5_append_angular_fluence: particle group");
    hid_t a = open_group(p, uncollided ? "uncollided_angular_fluence" : "angular_fluence");
    hid_t eg = a >= 0 ? open_group(a, "energy_" + std::to_string(g)) : -1;
    int rc = eg < 0 ? fail("lbh5_append_angular_fluence: groups") : 0;
    const hsize_t d[1] = {(hsize_t)NENK};
    for (int64_t i = 0; i < NI && !rc; ++i)
      rc = dset_f64(eg, "angle_" + std::to_string(i + 1), 1, d, v + i * NENK);
    if (eg >= 0) H5Gclose(eg);
    if (a >= 0) H5Gclose(a);
    H5Gclose(p);
    return rc;
  });
}

int lbh5_append_fluence(lbh5* h, const char* particle, int uncollided, int g, int64_t NENK, const double* phi) {
  return guarded([&]() -> int {
    if (!h || !particle || !phi || g < 1 || NENK <= 0) return fail("lbh5_append_fluence: invalid argument");
    hid_t p = open_group(h->file, lower(particle));
    if (p < 0) return fail("lbh5_append_fluence: particle group");
    hid_t f = open_group(p, uncollided ? "uncollided_fluence" : "fluence");
    const hsize_t d[1] = {(hsize_t)NENK};
    int rc = f < 0 ? fail("lbh5_append_fluence: group") : dset_f64(f, "energy_" + std::to_string(g), 1, d, phi);
    if (f >= 0) H5Gclose(f);
    H5Gclose(p);
    return rc;
  });
}

int lbh5_close(lbh5* h) {
  return guarded([&]() -> int {
    if (!h) return 0;
    herr_t e = h->file >= 0 ? H5Fclose(h->file) : 0;
    delete h;
    return e < 0 ? fail("lbh5_close: H5Fclose failed") : 0;
  });
}

}  // extern "C"
