"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
use this module, and only as the checker or the timed CPU reference.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i32 = ctypes.POINTER(ctypes.c_int32)
_ci = ctypes.POINTER(ctypes.c_int)


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_op_create.restype = ctypes.c_void_p
        L.orc_op_create.argtypes = [ctypes.c_int64, ctypes.c_int64, _d, _i64, _i64, _i64,
                                    ctypes.c_int64, _d, _d, ctypes.c_int]
        L.orc_op_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_physics.argtypes = [ctypes.c_void_p, _d, _d]
        L.orc_get_geometry.argtypes = [ctypes.c_void_p] + [_d] * 6
        L.orc_get_schedule.argtypes = [ctypes.c_void_p, _i32, _i32, _i32]
        L.orc_get_yy.argtypes = [ctypes.c_void_p, _d]
        for name in ("orc_transport_matinv", "orc_sweep_only", "orc_scatter_matvec", "orc_boltzmann_matvec"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _d]
        L.orc_gmres.restype = ctypes.c_int
        L.orc_gmres.argtypes = [ctypes.c_void_p, ctypes.c_int, _d, _d, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int, _d, ctypes.c_int, _ci, _ci]
        L.orc_source_iteration.restype = ctypes.c_int
        L.orc_source_iteration.argtypes = [ctypes.c_void_p, _d, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, _d, ctypes.c_int, _ci]
        L.orc_boltzmann_matinv.restype = ctypes.c_int
        L.orc_boltzmann_matinv.argtypes = [ctypes.c_void_p, _d, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, _d, ctypes.c_int, _ci, _ci]
        L.orc_build_delta_down.argtypes = [ctypes.c_void_p, _d]
        L.orc_scatter_moment_matvec.argtypes = [ctypes.c_void_p, _d, _d, _i32, ctypes.c_int, _d]
        L.orc_use_moment_kernel.argtypes = [ctypes.c_void_p, _d, _d, _i32, ctypes.c_int]
        L.orc_fluence.argtypes = [ctypes.c_void_p, _d, _d]
        L.orc_beam_uncollided.restype = ctypes.c_int
        L.orc_beam_uncollided.argtypes = [ctypes.c_void_p, _d, _d, _d, _d, ctypes.c_int, _d, ctypes.c_double, _d, _d,
                                          _i32]
        L.orc_beam_source.argtypes = [ctypes.c_void_p, _d, _d, _d]
        L.orc_beam_angular_fluence.argtypes = [ctypes.c_void_p, _d, _d, _d]
        L.orc_norm2.restype = ctypes.c_double
        L.orc_norm2.argtypes = [_d, ctypes.c_int64]
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


class OracleOp:
    """The reference operator stack (TransportOp + ScatteringOp + BoltzmannOp) on the CPU."""

    def __init__(self, mesh, ang):
        L = lib()
        self.mesh, self.ang = mesh, ang
        self.NE, self.NI = mesh.NE, ang.NI
        self.NENK = 4 * mesh.NE
        self.n = self.NENK * self.NI
        self._keep = [np.ascontiguousarray(a) for a in (mesh.rg, mesh.node, mesh.nbr, mesh.mat, ang.k, ang.w)]
        rg, node, nbr, mat, k, w = self._keep
        self.h = L.orc_op_create(mesh.NE, mesh.NK, _p(rg), _p(node, _i64), _p(nbr, _i64), _p(mat, _i64),
                                 ang.NI, _p(k), _p(w), ang.L)
        if not self.h:
            raise RuntimeError("SweepOrder.f90: Sweep is stagnant")
        self.Nm = int(mesh.mat.max())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_op_destroy(self.h)
            self.h = None

    def set_physics(self, sigt, sigs):
        sigt = np.ascontiguousarray(np.atleast_1d(sigt), dtype=np.float64)
        sigs = np.ascontiguousarray(np.atleast_2d(sigs), dtype=np.float64)
        assert sigt.shape == (self.Nm,) and sigs.shape == (self.Nm, self.ang.L + 1)
        lib().orc_set_physics(self.h, _p(sigt), _p(sigs))
        self.sigt = sigt

    def geometry(self):
        NE = self.NE
        IC = np.empty((NE, 3, 4, 4))
        IFM = np.empty((NE, 4, 4, 4))
        IIM = np.empty((NE, 4, 4, 4))
        nrm = np.empty((NE, 4, 3))
        area = np.empty((NE, 4))
        detJ = np.empty(NE)
        lib().orc_get_geometry(self.h, _p(IC), _p(IFM), _p(IIM), _p(nrm), _p(area), _p(detJ))
        return dict(IC=IC, IFM=IFM, IIM=IIM, nrm=nrm, area=area, detJ=detJ)

    def schedule(self):
        sl = np.empty((self.NI, self.NE), dtype=np.int32)
        nlev = np.empty(self.NI, dtype=np.int32)
        lev = np.empty((self.NI, self.NE), dtype=np.int32)
        lib().orc_get_schedule(self.h, _p(sl, _i32), _p(nlev, _i32), _p(lev, _i32))
        return sl, nlev, lev

    def yy(self):
        out = np.empty((self.ang.L + 1, self.NI, self.NI))
        lib().orc_get_yy(self.h, _p(out))
        return out

    def _vec(self, x):
        x = np.array(x, dtype=np.float64, copy=True).reshape(-1)
        assert x.size == self.n
        return x

    def transport_matinv(self, s):
        x = self._vec(s)
        lib().orc_transport_matinv(self.h, _p(x))
        return x

    def sweep_only(self, s):
        x = self._vec(s)
        lib().orc_sweep_only(self.h, _p(x))
        return x

    def scatter_matvec(self, s):
        x = self._vec(s)
        lib().orc_scatter_matvec(self.h, _p(x))
        return x

    def build_delta_down(self, xsdd):
        """ScatteringOp%BuildDeltaDown: K = XSdd(mat) I until the next set_physics."""
        self._xsdd = np.ascontiguousarray(xsdd, dtype=np.float64)
        assert self._xsdd.size >= self.Nm
        lib().orc_build_delta_down(self.h, _p(self._xsdd))

    def scatter_moment_matvec(self, s, D, M, deg):
        """K x through the flux moments (the CPU baseline's moment form)."""
        x = self._vec(s)
        D = np.ascontiguousarray(D, dtype=np.float64)
        M = np.ascontiguousarray(M, dtype=np.float64)
        deg = np.ascontiguousarray(deg, dtype=np.int32)
        assert D.shape == (deg.size, self.NI) and M.shape == (self.NI, deg.size)
        lib().orc_scatter_moment_matvec(self.h, _p(D), _p(M), deg.ctypes.data_as(_i32), deg.size, _p(x))
        return x

    def use_moment_kernel(self, D=None, M=None, deg=None):
        """Apply K in moment form from now on (every scatter/Boltzmann product and
        solve); no arguments restores the dense kernel.  For full-size checks, where
        the dense kernel's 2 NI^2 flops per dof would take minutes per product."""
        if D is None:
            lib().orc_use_moment_kernel(self.h, None, None, None, 0)
            return
        self._kD = np.ascontiguousarray(D, dtype=np.float64)
        self._kM = np.ascontiguousarray(M, dtype=np.float64)
        self._kdeg = np.ascontiguousarray(deg, dtype=np.int32)
        assert self._kD.shape == (self._kdeg.size, self.NI) and self._kM.shape == (self.NI, self._kdeg.size)
        lib().orc_use_moment_kernel(self.h, _p(self._kD), _p(self._kM), self._kdeg.ctypes.data_as(_i32),
                                    self._kdeg.size)

    def boltzmann_matvec(self, s):
        x = self._vec(s)
        lib().orc_boltzmann_matvec(self.h, _p(x))
        return x

    def boltzmann_matinv(self, s, kind="gmres", m=10, rtol=1e-8, atol=1e-12, maxit=0, hist_cap=100000):
        x = self._vec(s)
        hist = np.zeros(hist_cap)
        iters = ctypes.c_int(0)
        nmv = ctypes.c_int(0)
        st = lib().orc_boltzmann_matinv(self.h, _p(x), 0 if kind == "gmres" else 1, m, rtol, atol, maxit,
                                        _p(hist), hist_cap, ctypes.byref(iters), ctypes.byref(nmv))
        return x, dict(status=st, iters=iters.value, nmatvec=nmv.value, hist=hist[:iters.value].copy())

    def gmres(self, rhs, x0, m=10, rtol=1e-8, atol=1e-12, maxit=0, hist_cap=100000):
        rhs = self._vec(rhs)
        x = self._vec(x0)
        hist = np.zeros(hist_cap)
        iters = ctypes.c_int(0)
        nmv = ctypes.c_int(0)
        st = lib().orc_gmres(self.h, m, _p(rhs), _p(x), rtol, atol, maxit, _p(hist), hist_cap,
                             ctypes.byref(iters), ctypes.byref(nmv))
        return x, dict(status=st, iters=iters.value, nmatvec=nmv.value, hist=hist[:iters.value].copy())

    def fluence(self, v):
        v = self._vec(v)
        out = np.empty(self.NENK)
        lib().orc_fluence(self.h, _p(v), _p(out))
        return out

    # -- planar external beam (the reference's ExternalBeam, planar branches) ----
    @staticmethod
    def beam_frame(axis):
        """An orthonormal pair spanning the plane normal to the beam axis."""
        a = np.asarray(axis, dtype=np.float64)
        t = np.array([1.0, 0.0, 0.0]) if abs(a[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e1 = t - a.dot(t) * a
        e1 /= np.linalg.norm(e1)
        return e1, np.cross(a, e1)

    def beam_uncollided(self, axis, origin, attn, cutout="none", params=None, n=1.0, frame=None):
        """phi0ek (NENK,): the projected uncollided fluence, and the in-beam mask of
        the quadrature points.  ``attn``: sigma_t per material."""
        axis = np.ascontiguousarray(axis, dtype=np.float64)
        origin = np.ascontiguousarray(origin, dtype=np.float64)
        e1, e2 = self.beam_frame(axis) if frame is None else frame
        e1, e2 = np.ascontiguousarray(e1, dtype=np.float64), np.ascontiguousarray(e2, dtype=np.float64)
        prm = np.zeros(4) if params is None else np.ascontiguousarray(params, dtype=np.float64)
        attn = np.ascontiguousarray(np.atleast_1d(attn), dtype=np.float64)
        assert attn.size == self.Nm
        out = np.empty(self.NENK)
        mask = np.empty(self.NENK, dtype=np.int32)
        cut = {"none": 0, "circle": 1, "rectangle": 2}[cutout]
        rc = lib().orc_beam_uncollided(self.h, _p(axis), _p(origin), _p(e1), _p(e2), cut, _p(prm), float(n),
                                       _p(attn), _p(out), _p(mask, _i32))
        if rc != 0:
            raise RuntimeError("RayTracing.f90: TraceVector: While ray tracing no intersections were found.")
        return out, mask.astype(bool)

    def beam_source(self, axis, phi0ek, s):
        """FCSSource (planar): s + phi0ek x sum_l (2l+1)/(4 pi) P_l(k.k0) sigma_l(mat)."""
        x = self._vec(s)
        axis = np.ascontiguousarray(axis, dtype=np.float64)
        phi0ek = np.ascontiguousarray(phi0ek, dtype=np.float64)
        lib().orc_beam_source(self.h, _p(axis), _p(phi0ek), _p(x))
        return x

    def beam_angular_fluence(self, axis, phi0ek):
        """AddAngularFluence (planar): the uncollided angular fluence."""
        unc = np.zeros(self.n)
        axis = np.ascontiguousarray(axis, dtype=np.float64)
        phi0ek = np.ascontiguousarray(phi0ek, dtype=np.float64)
        lib().orc_beam_angular_fluence(self.h, _p(axis), _p(phi0ek), _p(unc))
        return unc


def norm2(x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    return lib().orc_norm2(_p(x), x.size)


def num_threads():
    return lib().orc_num_threads()
