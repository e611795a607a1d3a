"""Results in the reference's HDF5 layout (liblbh5, include/lionbolt_h5.h).

``write_results`` produces the file Main.f90 produces for one particle: the
problem type, the mesh group (WriteMeshH5), the particle group with its
ordinates and energy grid, and per energy group the angular fluence
(``angular_fluence/energy_<g>/angle_<i>``) and, when given, the scalar fluence
(``fluence/energy_<g>``) and their uncollided counterparts
(Main.f90:436-470, LionboltHDF5.f90:15-344).
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblbh5.so")
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
SIGNATURES = {
    "lbh5_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "lbh5_write_problem_type": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "lbh5_write_mesh": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _i64, _d, _i64, _d, ctypes.c_int]),
    "lbh5_write_particle": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int64, _d, _d, ctypes.c_int, _d]),
    "lbh5_append_angular_fluence": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                                   ctypes.c_int64, _d]),
    "lbh5_append_fluence": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _d]),
    "lbh5_close": (ctypes.c_int, [_vp]),
    "lbh5_last_error": (ctypes.c_char_p, []),
}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (make -C lionbolt_amd/csrc needs the HDF5 C library)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise IOError(lib().lbh5_last_error().decode())


def _d64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_d)


def write_results(path, mesh, ang, particle="photons", psi=None, phi=None, psi_unc=None, phi_unc=None,
                  energy_grid=None, problem_type="general"):
    """Write one particle's results.

    ``psi``: list over groups of angular fluences in the reference layout
    (NI * NENK, direction-major); ``phi``: list over groups of scalar fluences
    (NENK); ``psi_unc`` / ``phi_unc``: the uncollided ones.  Groups are numbered
    from 1 in the file, as the reference numbers them."""
    L = lib()
    h = _vp()
    _check(L.lbh5_create(os.fsencode(path), ctypes.byref(h)))
    keep = []
    try:
        _check(L.lbh5_write_problem_type(h, problem_type.encode()))
        node = np.ascontiguousarray(mesh.node, dtype=np.int64)
        mat = np.ascontiguousarray(mesh.mat, dtype=np.int64)
        rg, rgp = _d64(mesh.rg)
        vol, volp = _d64(mesh.volumes())
        keep += [node, mat, rg, vol]
        _check(L.lbh5_write_mesh(h, mesh.NE, mesh.NK, node.ctypes.data_as(_i64), rgp, mat.ctypes.data_as(_i64),
                                 volp, int(problem_type == "slab")))
        w, wp = _d64(ang.w)
        k, kp = _d64(ang.k)
        G = 0 if energy_grid is None else len(energy_grid) - 1
        E, Ep = _d64(energy_grid if energy_grid is not None else np.zeros(1))
        keep += [w, k, E]
        _check(L.lbh5_write_particle(h, particle.encode(), ang.NI, wp, kp, G, Ep if G > 0 else None))
        nenk = 4 * mesh.NE
        for unc, series in ((0, psi), (1, psi_unc)):
            for g, v in enumerate(series or [], start=1):
                v, vp = _d64(np.asarray(v).reshape(-1))
                if v.size != ang.NI * nenk:
                    raise ValueError("angular fluence must hold NI * NENK values")
                _check(L.lbh5_append_angular_fluence(h, particle.encode(), unc, g, ang.NI, nenk, vp))
        for unc, series in ((0, phi), (1, phi_unc)):
            for g, f in enumerate(series or [], start=1):
                f, fp = _d64(np.asarray(f).reshape(-1))
                if f.size != nenk:
                    raise ValueError("fluence must hold NENK values")
                _check(L.lbh5_append_fluence(h, particle.encode(), unc, g, nenk, fp))
    finally:
        _check(L.lbh5_close(h))
