"""Host-side mirror of the reference's operator API on top of liblbhip.

The reference exposes (src/LionboltAPI.f90:1-11, LinearOperatorClass.f90:30-112):

* ``SpaceAngleVector(mesh, angular)`` with ``%v`` and ``%Fluence(w)``;
* ``TransportOp``, ``ScatteringOp``, ``BoltzmannOp`` with the generic
  ``%Build(mesh, angular, XS)`` / ``%Build(mesh)`` / ``%Build(mesh, angular)`` /
  ``%Build(angular, XS)``, ``%MatVec(x)``, ``%MatInv(s)`` and ``%Destroy()``.

The classes below keep those names, argument meanings and in-place semantics;
every failure raises ``LionboltError`` where the reference calls ``stophere``.
``Operator`` is the thin handle wrapper the classes share, which also exposes
the device-resident entry points.
"""
import ctypes
import sys

import numpy as np

from . import _lib
from .angular import AngularSet
from .mesh import TetMesh

MeshClass = TetMesh
AngularClass = AngularSet


def _torch_done():
    """The library works on its own non-blocking stream: wait for the caller's
    current torch stream so that tensors it just produced are complete."""
    import torch
    torch.cuda.current_stream().synchronize()


def _dev(t, need, name, dtype=None):
    """Device pointer of a tensor handed to a ``lb_*_dev`` entry point, after
    checking what the library assumes of it: float64 (or ``dtype``), on the GPU,
    contiguous and holding at least ``need`` entries.  A tensor that fails any
    check would let the kernels read or write past its end."""
    import torch
    dtype = torch.float64 if dtype is None else dtype
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: expected {dtype}, got {t.dtype}")
    if not t.is_cuda:
        raise ValueError(f"{name}: expected a tensor on the GPU")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    if t.numel() < need:
        raise ValueError(f"{name}: {t.numel()} entries, at least {need} needed")
    return t.data_ptr()


class XSType:
    """Cross sections of one group: ``t(Nm)`` and Legendre moments ``s(0:L, Nm)``."""

    def __init__(self, t, s):
        self.t = np.atleast_1d(np.asarray(t, dtype=np.float64))
        self.s = np.asarray(s, dtype=np.float64).reshape(-1, self.t.size)   # (L+1, Nm)


class Operator:
    """One liblbhip operator handle (geometry + schedule + physics)."""

    def __init__(self, mesh, ang):
        L = _lib.lib()
        self.mesh, self.ang = mesh, ang
        self.NE, self.NI = mesh.NE, ang.NI
        self.NENK = 4 * mesh.NE
        self.n = self.NENK * self.NI
        self.nmom = (ang.L + 1) ** 2
        self._keep = [np.ascontiguousarray(a) for a in (mesh.rg, mesh.node, mesh.nbr, mesh.mat - 1, ang.k, ang.w)]
        rg, node, nbr, mat, k, w = self._keep
        self._mesh = _lib.MeshDesc(mesh.NE, mesh.NK, _lib.dptr(rg), _lib.i64ptr(node), _lib.i64ptr(nbr),
                                   _lib.i64ptr(mat), 0)
        self._quad = _lib.QuadDesc(ang.NI, _lib.dptr(k), _lib.dptr(w), ang.L)
        h = ctypes.c_void_p()
        _lib.check(L.lb_op_create(ctypes.byref(self._mesh), ctypes.byref(self._quad), ctypes.byref(h)))
        self.h = h
        ne, ni, nm, ml = (ctypes.c_int64() for _ in range(4))
        _lib.check(L.lb_op_info(h, ctypes.byref(ne), ctypes.byref(ni), ctypes.byref(nm), ctypes.byref(ml)))
        self.maxlev = ml.value
        self.Nm = max(1, int(np.max(mesh.mat)))    # materials 1..Nm, as the library counts them
        self.G = 0

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().lb_op_destroy(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the ctypes binding and the HIP runtime may already be
        # gone; the process releases the device memory itself
        if not sys.is_finalizing():
            self.close()

    # -- physics ---------------------------------------------------------------
    def set_physics(self, sigt, sigs):
        """One group: ``sigt`` (Nm,), ``sigs`` (Nm, L+1)."""
        sigt = np.ascontiguousarray(np.atleast_1d(sigt), dtype=np.float64)
        sigs = np.ascontiguousarray(np.atleast_2d(sigs), dtype=np.float64)
        _lib.check(_lib.lib().lb_op_set_physics(self.h, sigt.size, _lib.dptr(sigt), sigs.shape[1] - 1,
                                                _lib.dptr(sigs)))
        self.G = 1
        self.sigt = sigt.reshape(1, -1)

    def build_delta_down(self, xsdd):
        """ScatteringOp%BuildDeltaDown: K = XSdd(mat) I until the next set_physics."""
        xsdd = np.ascontiguousarray(np.atleast_1d(xsdd), dtype=np.float64)
        _lib.check(_lib.lib().lb_op_build_delta_down(self.h, xsdd.size, _lib.dptr(xsdd)))

    def set_physics_groups(self, sigt, sigs):
        """``sigt`` (G, Nm), ``sigs`` (G, Nm, L+1)."""
        sigt = np.ascontiguousarray(sigt, dtype=np.float64)
        sigs = np.ascontiguousarray(sigs, dtype=np.float64)
        G, Nm = sigt.shape
        _lib.check(_lib.lib().lb_op_set_physics_groups(self.h, G, Nm, _lib.dptr(sigt), sigs.shape[2] - 1,
                                                       _lib.dptr(sigs)))
        self.G = G
        self.sigt = sigt

    def set_scatter_mode(self, mode):
        m = {"moment": _lib.SCATTER_MOMENT, "dense": _lib.SCATTER_DENSE}[mode]
        _lib.check(_lib.lib().lb_op_set_scatter_mode(self.h, m))

    _SWEEP_MODES = {"levels": _lib.SWEEP_LEVELS, "graph": _lib.SWEEP_GRAPH, "direction": _lib.SWEEP_DIRECTION,
                    "auto": _lib.SWEEP_AUTO, "persistent": _lib.SWEEP_PERSISTENT,
                    "direction_lds": _lib.SWEEP_DIRECTION_LDS}

    def set_sweep_mode(self, mode):
        _lib.check(_lib.lib().lb_op_set_sweep_mode(self.h, self._SWEEP_MODES[mode]))

    def sweep_choice(self, G=1):
        """(mode name, workgroup size, persistent workgroups per CU) the autotuner settled
        on for a group batch of G, or ("auto", 0, 0) while it is still timing."""
        mode, threads, per_cu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().lb_op_sweep_choice(self.h, int(G), ctypes.byref(mode), ctypes.byref(threads),
                                                 ctypes.byref(per_cu)))
        name = {v: k for k, v in self._SWEEP_MODES.items()}[mode.value]
        return name, threads.value, per_cu.value

    def sweep_order_choice(self, source_in_order=False):
        """(mode name, workgroup size) the autotuner settled on for one-group sweeps that
        write a sweep-order flux (from a natural or a sweep-order source), or ("auto", 0)
        while it is still timing."""
        mode, threads = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().lb_op_sweep_order_choice(self.h, int(bool(source_in_order)), ctypes.byref(mode),
                                                       ctypes.byref(threads)))
        name = {v: k for k, v in self._SWEEP_MODES.items()}[mode.value]
        return name, threads.value

    def set_flux_order(self, order):
        """Layout of one group's angular flux inside the within-group flux-moment solves:
        "natural", "sweep" (each direction's flux stored in its sweep order),
        "sweep_source" (the sweep's source too), which all give bit-identical
        solutions, or "sweep_fp32" (the Arnoldi steps' flux for D stored in FP32)."""
        codes = {"natural": _lib.FLUX_NATURAL, "sweep": _lib.FLUX_SWEEP_ORDER,
                 "sweep_source": _lib.FLUX_SWEEP_ORDER_SOURCE, "sweep_fp32": _lib.FLUX_SWEEP_ORDER_FP32}
        if order not in codes:
            raise ValueError(f"flux order must be one of {sorted(codes)}")
        _lib.check(_lib.lib().lb_op_set_flux_order(self.h, codes[order]))

    def set_group_batch(self, groups):
        """Energy groups solved together per coupled system in ``multigroup_solve_dev``."""
        _lib.check(_lib.lib().lb_op_set_group_batch(self.h, int(groups)))

    def set_workgroups_per_cu(self, per_cu):
        _lib.check(_lib.lib().lb_op_set_workgroups_per_cu(self.h, int(per_cu)))

    def set_problems(self, elem_off):
        """Declare the mesh as independent problems in element ranges [elem_off[p], elem_off[p+1])."""
        off = np.ascontiguousarray(elem_off, dtype=np.int64)
        _lib.check(_lib.lib().lb_op_set_problems(self.h, off.size - 1, _lib.i64ptr(off)))
        self.P = off.size - 1

    def batch_solve_dev(self, qmom, phi, m=30, rtol=1e-8, atol=1e-12, maxit=0, raise_on_fail=True):
        """Per-problem flux-moment GMRES(m) of the whole batch in lockstep (single group).
        ``qmom``/``phi``: cuda float64 tensors of nmom * NENK entries."""
        nphi = self.nmom * self.NENK
        q_p, phi_p = _dev(qmom, nphi, "qmom"), _dev(phi, nphi, "phi")
        _torch_done()
        opts = _lib.SolverOpts(_lib.SOLVER_GMRES, _lib.SPACE_MOMENT, m, maxit, rtol, atol, 0)
        stats = (_lib.SolveStats * self.P)()
        rc = _lib.lib().lb_batch_solve_dev(self.h, q_p, ctypes.byref(opts), phi_p, stats)
        if raise_on_fail:
            _lib.check(rc)
        return [s.as_dict() for s in stats]

    def set_direction_block(self, threads):
        _lib.check(_lib.lib().lb_op_set_direction_block(self.h, int(threads)))

    # -- host entry points (reference layout v[sdof + i*NENK]) -------------------
    def _vec(self, x):
        x = np.array(x, dtype=np.float64, copy=True).reshape(-1)
        if x.size != self.n:
            raise ValueError(f"vector length {x.size} != NENK*NI = {self.n}")
        return x

    def transport_matinv(self, s):
        x = self._vec(s)
        _lib.check(_lib.lib().lb_transport_matinv(self.h, _lib.dptr(x), x.size))
        return x

    def scatter_matvec(self, s):
        x = self._vec(s)
        _lib.check(_lib.lib().lb_scatter_matvec(self.h, _lib.dptr(x), x.size))
        return x

    def boltzmann_matvec(self, s):
        x = self._vec(s)
        _lib.check(_lib.lib().lb_boltzmann_matvec(self.h, _lib.dptr(x), x.size))
        return x

    def boltzmann_matinv(self, s, kind="gmres", space="psi", m=10, rtol=1e-8, atol=1e-12, maxit=0, verbose=False,
                         raise_on_fail=True):
        x = self._vec(s)
        opts = _lib.SolverOpts({"gmres": _lib.SOLVER_GMRES, "si": _lib.SOLVER_SI, "fgmres": _lib.SOLVER_FGMRES}[kind],
                               {"psi": _lib.SPACE_PSI, "moment": _lib.SPACE_MOMENT}[space], m, maxit, rtol, atol,
                               int(verbose))
        st = _lib.SolveStats()
        rc = _lib.lib().lb_boltzmann_matinv(self.h, _lib.dptr(x), x.size, ctypes.byref(opts), ctypes.byref(st))
        if raise_on_fail:
            _lib.check(rc)
        return x, st.as_dict()

    def fluence(self, v):
        v = self._vec(v)
        out = np.empty(self.NENK)
        _lib.check(_lib.lib().lb_fluence(self.h, _lib.dptr(v), _lib.dptr(out)))
        return out

    def schedule(self):
        sl = np.empty((self.NI, self.NE), dtype=np.int32)
        nlev = np.empty(self.NI, dtype=np.int32)
        _lib.check(_lib.lib().lb_op_get_schedule(self.h, _lib.i32ptr(sl), _lib.i32ptr(nlev)))
        return sl, nlev

    def normals(self):
        n = np.empty((self.NE, 4, 3))
        _lib.check(_lib.lib().lb_op_get_normals(self.h, _lib.dptr(n)))
        return n

    # -- device entry points (torch tensors on cuda, group-batched layout) --------
    def transport_matinv_dev(self, q, psi, g0=0, G=1):
        """psi = T^-1 q over groups g0 .. g0 + G - 1 (layout [i][e][k][g]).  Asynchronous:
        the sweep is enqueued on the operator's stream and may still run on return;
        ``synchronize()`` waits for it (as the other device calls do before returning)."""
        need = self.n * int(G)
        q_p, psi_p = _dev(q, need, "q"), _dev(psi, need, "psi")
        _torch_done()
        _lib.check(_lib.lib().lb_transport_matinv_dev(self.h, q_p, psi_p, g0, G))

    def transport_matinv_order_dev(self, q, psi_so, g0=0, source_in_order=False):
        """One group's psi = T^-1 q with psi_so in sweep order: psi_so[(i NE + pos) 4 + k]
        is node k of the cell at position pos of direction i's device sweep order.
        q is in the natural layout (then it must not overlap psi_so) or, with
        ``source_in_order``, in sweep order too (then it may be psi_so itself).
        Asynchronous, as ``transport_matinv_dev``."""
        q_p, psi_p = _dev(q, self.n, "q"), _dev(psi_so, self.n, "psi_so")
        _torch_done()
        _lib.check(_lib.lib().lb_transport_matinv_order_dev(self.h, q_p, psi_p, int(g0), int(bool(source_in_order))))

    def moments_to_discrete_order_dev(self, phi, psi_so, g0=0):
        """One group's M S_g0 phi written in sweep order."""
        phi_p = _dev(phi, self.nmom * self.NENK, "phi")
        psi_p = _dev(psi_so, self.n, "psi_so")
        _torch_done()
        _lib.check(_lib.lib().lb_moments_to_discrete_order_dev(self.h, phi_p, psi_p, int(g0)))
        self.synchronize()

    def transport_matinv_order32_dev(self, q, psi_so, psi32_so, g0=0):
        """One group's T^-1 q with the flux in sweep order as FP32 in psi32_so (float32);
        psi_so (float64) receives only the cells whose flux is read from global memory."""
        import torch
        q_p, so_p = _dev(q, self.n, "q"), _dev(psi_so, self.n, "psi_so")
        f_p = _dev(psi32_so, self.n, "psi32_so", torch.float32)
        _torch_done()
        _lib.check(_lib.lib().lb_transport_matinv_order32_dev(self.h, q_p, so_p, f_p, int(g0)))

    def moments_order32_dev(self, psi32_so, phi):
        """phi = D psi for one group's FP32 sweep-order flux."""
        import torch
        f_p = _dev(psi32_so, self.n, "psi32_so", torch.float32)
        phi_p = _dev(phi, self.nmom * self.NENK, "phi")
        _torch_done()
        _lib.check(_lib.lib().lb_moments_order32_dev(self.h, f_p, phi_p))
        self.synchronize()

    def moments_order_dev(self, psi_so, phi):
        """phi = D psi for one group's sweep-order flux."""
        psi_p = _dev(psi_so, self.n, "psi_so")
        phi_p = _dev(phi, self.nmom * self.NENK, "phi")
        _torch_done()
        _lib.check(_lib.lib().lb_moments_order_dev(self.h, psi_p, phi_p))
        self.synchronize()

    def flux_to_natural_dev(self, psi_so, psi):
        """One group's sweep-order flux in the natural layout [i][e][k]."""
        so_p, psi_p = _dev(psi_so, self.n, "psi_so"), _dev(psi, self.n, "psi")
        _torch_done()
        _lib.check(_lib.lib().lb_flux_to_natural_dev(self.h, so_p, psi_p))
        self.synchronize()

    def moments_dev(self, psi, phi, G=1):
        psi_p = _dev(psi, self.n * int(G), "psi")
        phi_p = _dev(phi, self.nmom * self.NENK * int(G), "phi")
        _torch_done()
        _lib.check(_lib.lib().lb_moments_dev(self.h, psi_p, phi_p, G))
        self.synchronize()

    def moments_to_discrete_dev(self, phi, psi, g0=0, G=1, accumulate=False):
        phi_p = _dev(phi, self.nmom * self.NENK * int(G), "phi")
        psi_p = _dev(psi, self.n * int(G), "psi")
        _torch_done()
        _lib.check(_lib.lib().lb_moments_to_discrete_dev(self.h, phi_p, psi_p, g0, G, int(accumulate)))
        self.synchronize()

    def multigroup_solve_dev(self, transfer, phi, qmom=None, kind="gmres", space="moment", m=30, rtol=1e-8,
                             atol=1e-12, maxit=0, raise_on_fail=True):
        """Group-by-group downscatter solve; ``transfer`` (G, G, Nm, L+1) host array [g', g]."""
        need = self.G * self.nmom * self.NENK
        phi_p = _dev(phi, need, "phi")
        q_p = None if qmom is None else _dev(qmom, need, "qmom")
        _torch_done()
        transfer = np.ascontiguousarray(transfer, dtype=np.float64)
        opts = _lib.SolverOpts({"gmres": _lib.SOLVER_GMRES, "si": _lib.SOLVER_SI, "fgmres": _lib.SOLVER_FGMRES}[kind],
                               {"psi": _lib.SPACE_PSI, "moment": _lib.SPACE_MOMENT}[space], m, maxit, rtol, atol, 0)
        stats = (_lib.SolveStats * self.G)()
        rc = _lib.lib().lb_multigroup_solve_dev(self.h, q_p, _lib.dptr(transfer), ctypes.byref(opts), phi_p, stats)
        if raise_on_fail:
            _lib.check(rc)
        return [s.as_dict() for s in stats]

    # -- learned right preconditioner and the within-group moment system ---------
    def set_preconditioner(self, params, fmt="bf16"):
        """Per-cell-node MLP parameters (``_lib.MLP_PARAMS`` float32) for FGMRES."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        f = {"bf16": _lib.FMT_BF16, "fp32": _lib.FMT_FP32}[fmt]
        _lib.check(_lib.lib().lb_op_set_preconditioner(self.h, p.ctypes.data_as(_lib._f), p.size, f))

    def set_preconditioner_threshold(self, min_thickness):
        """Groups optically thinner than ``min_thickness`` skip the network (plain GMRES)."""
        _lib.check(_lib.lib().lb_op_set_preconditioner_threshold(self.h, float(min_thickness)))

    _MOMENT_FMTS = {"fp64": _lib.FMT_FP64, "fp32": _lib.FMT_FP32, "bf16": _lib.FMT_BF16,
                    "split_bf16": _lib.FMT_SPLIT_BF16, "fp16": _lib.FMT_FP16}

    def set_moment_precision(self, fmt="fp64"):
        """Input precision of D psi inside the flux-moment GMRES's Arnoldi steps
        (MFMA formats with FP32 accumulation; the right-hand side, psi and the true
        residual stay FP64)."""
        _lib.check(_lib.lib().lb_op_set_moment_precision(self.h, self._MOMENT_FMTS[fmt]))

    def mlp_apply_dev(self, r, z, g0=0):
        nphi = self.nmom * self.NENK
        r_p, z_p = _dev(r, nphi, "r"), _dev(z, nphi, "z")
        _torch_done()
        _lib.check(_lib.lib().lb_mlp_apply_dev(self.h, r_p, z_p, int(g0)))
        self.synchronize()

    def mlp_apply_noalloc_dev(self, r, z, g0=0):
        """``mlp_apply_dev`` through the entry point that allocates nothing (its
        reduction scratch is a per-GPU device array; overlapping calls on one GPU race)."""
        nphi = self.nmom * self.NENK
        r_p, z_p = _dev(r, nphi, "r"), _dev(z, nphi, "z")
        _torch_done()
        _lib.check(_lib.lib().lb_mlp_apply_noalloc_dev(self.h, r_p, z_p, int(g0)))
        self.synchronize()

    def mlp_features_dev(self, r, g0=0):
        r_p = _dev(r, self.nmom * self.NENK, "r")
        _torch_done()
        import torch
        feat = torch.empty((self.NENK, 18), dtype=torch.float32, device=r.device)
        s = ctypes.c_double()
        _lib.check(_lib.lib().lb_mlp_features_dev(self.h, r_p, int(g0), feat.data_ptr(), ctypes.byref(s)))
        self.synchronize()
        return feat, s.value

    def moment_rhs_dev(self, q, b, g0=0):
        q_p, b_p = _dev(q, self.n, "q"), _dev(b, self.nmom * self.NENK, "b")
        _torch_done()
        _lib.check(_lib.lib().lb_moment_rhs_dev(self.h, q_p, b_p, int(g0)))
        self.synchronize()

    def discrete_source_dev(self, qmom, q):
        qm_p, q_p = _dev(qmom, self.nmom * self.NENK, "qmom"), _dev(q, self.n, "q")
        _torch_done()
        _lib.check(_lib.lib().lb_discrete_source_dev(self.h, qm_p, q_p))
        self.synchronize()

    def moment_matvec_dev(self, x, y, g0=0):
        nphi = self.nmom * self.NENK
        x_p, y_p = _dev(x, nphi, "x"), _dev(y, nphi, "y")
        _torch_done()
        _lib.check(_lib.lib().lb_moment_matvec_dev(self.h, x_p, y_p, int(g0)))
        self.synchronize()

    def gmres_snapshots_dev(self, q, phi, snaps, m=30, rtol=1e-10, atol=1e-12, maxit=0):
        """Flux-moment GMRES(m) recording the iterate after every Arnoldi step into
        ``snaps`` (max_snaps x nmom*NENK); returns (count, stats)."""
        nphi = self.nmom * self.NENK
        q_p, phi_p = _dev(q, self.n, "q"), _dev(phi, nphi, "phi")
        s_p = _dev(snaps, snaps.shape[0] * nphi, "snaps")
        _torch_done()
        opts = _lib.SolverOpts(_lib.SOLVER_GMRES, _lib.SPACE_MOMENT, m, maxit, rtol, atol, 0)
        st = _lib.SolveStats()
        cnt = ctypes.c_int()
        _lib.check(_lib.lib().lb_gmres_snapshots_dev(self.h, q_p, ctypes.byref(opts), phi_p, s_p, snaps.shape[0],
                                                     ctypes.byref(cnt), ctypes.byref(st)))
        return cnt.value, st.as_dict()

    def synchronize(self):
        _lib.check(_lib.lib().lb_op_synchronize(self.h))

    def last_sweep_ms(self):
        ms = ctypes.c_double()
        _lib.check(_lib.lib().lb_last_sweep_ms(self.h, ctypes.byref(ms)))
        return ms.value


def schedule_host(mesh, ang):
    """The upwind schedule and normals computed by the library's host half (no GPU needed)."""
    keep = [np.ascontiguousarray(a) for a in (mesh.rg, mesh.node, mesh.nbr, mesh.mat - 1, ang.k, ang.w)]
    rg, node, nbr, mat, k, w = keep
    md = _lib.MeshDesc(mesh.NE, mesh.NK, _lib.dptr(rg), _lib.i64ptr(node), _lib.i64ptr(nbr), _lib.i64ptr(mat), 0)
    qd = _lib.QuadDesc(ang.NI, _lib.dptr(k), _lib.dptr(w), ang.L)
    sl = np.empty((ang.NI, mesh.NE), dtype=np.int32)
    nlev = np.empty(ang.NI, dtype=np.int32)
    nrm = np.empty((mesh.NE, 4, 3))
    _lib.check(_lib.lib().lb_schedule_host(ctypes.byref(md), ctypes.byref(qd), _lib.i32ptr(sl), _lib.i32ptr(nlev),
                                           _lib.dptr(nrm)))
    return sl, nlev, nrm


def trace_plan_host(mesh, ang, G, lds_bytes):
    """The LDS slot plan of the direction-resident LDS sweep on the schedule of
    ``schedule_host`` (no GPU needed): a dict with uslot [NI, NE, 4], wslot [NI, NE],
    stage_off [NI, maxlev + 2], stage [n, 2] (LDS slot, element), ring, stage_slots,
    stats (ring, staged, one level back, overflow) and gprod [NI, NE] (1 = the
    cell's flux is read from global memory by some consumer)."""
    keep = [np.ascontiguousarray(a) for a in (mesh.rg, mesh.node, mesh.nbr, mesh.mat - 1, ang.k, ang.w)]
    rg, node, nbr, mat, k, w = keep
    md = _lib.MeshDesc(mesh.NE, mesh.NK, _lib.dptr(rg), _lib.i64ptr(node), _lib.i64ptr(nbr), _lib.i64ptr(mat), 0)
    qd = _lib.QuadDesc(ang.NI, _lib.dptr(k), _lib.dptr(w), ang.L)
    _, nlev, _ = schedule_host(mesh, ang)
    maxlev = int(nlev.max())
    uslot = np.empty((ang.NI, mesh.NE, 4), dtype=np.int32)
    wslot = np.empty((ang.NI, mesh.NE), dtype=np.int32)
    soff = np.empty((ang.NI, maxlev + 2), dtype=np.int32)
    sizes = np.empty(2, dtype=np.int32)
    stats = np.empty(4, dtype=np.int64)
    gprod = np.empty((ang.NI, mesh.NE), dtype=np.uint8)
    n = ctypes.c_int64(0)
    L = _lib.lib()
    _lib.check(L.lb_trace_plan_host(ctypes.byref(md), ctypes.byref(qd), G, lds_bytes, None, None, None, None,
                                    ctypes.byref(n), None, None, None))
    stage = np.empty((max(n.value, 1), 2), dtype=np.int32)
    _lib.check(L.lb_trace_plan_host(ctypes.byref(md), ctypes.byref(qd), G, lds_bytes, _lib.i32ptr(uslot),
                                    _lib.i32ptr(wslot), _lib.i32ptr(soff), _lib.i32ptr(stage), ctypes.byref(n),
                                    _lib.i32ptr(sizes), stats.ctypes.data_as(_lib._i64), gprod.ctypes.data))
    return dict(uslot=uslot, wslot=wslot, stage_off=soff, stage=stage[:n.value], ring=int(sizes[0]),
                stage_slots=int(sizes[1]), stats=stats, nlev=nlev, gprod=gprod)


class SpaceAngleVector:
    """``SpaceAngleVector(mesh, angular)``: flat FP64 ``v`` of length NENK*NI."""

    def __init__(self, mesh, angular):
        self.NENK = 4 * mesh.NE
        self.NI = angular.NI
        self.v = np.zeros(self.NENK * self.NI)

    def Fluence(self, w):
        """``phi(sdof) = sum_i w_i v(sdof + i NENK)`` (LinearOperatorClass.f90:154-174)."""
        return (np.asarray(w)[:, None] * self.v.reshape(self.NI, self.NENK)).sum(axis=0)

    def Destroy(self):
        self.v = np.zeros(0)


class _LinearOperator:
    name = "Linear Operator"

    def __init__(self):
        self.SPACEREADY = self.ANGLEREADY = self.PHYSREADY = False
        self.mesh = self.angular = self.XS = None
        self._op = None

    def Build(self, *args):
        """Generic Build: (mesh, angular, XS) | (mesh) | (mesh, angular) | (angular, XS)."""
        if len(args) == 3:
            self.Build(args[0])
            self.Build(args[0], args[1])
            self.Build(args[1], args[2])
        elif len(args) == 1:
            self.mesh = args[0]
            self.SPACEREADY = True
        elif len(args) == 2 and isinstance(args[0], TetMesh):
            mesh, ang = args
            self.mesh, self.angular = mesh, ang
            self.SPACEREADY = True
            if self._op is not None:
                self._op.close()
            self._op = Operator(mesh, ang)
            self.ANGLEREADY = True
            if self.XS is not None:
                self._physics()
        elif len(args) == 2:
            self.XS = args[1]
            if self._op is not None:
                self._physics()
        else:
            raise TypeError("Build(mesh, angular, XS) | Build(mesh) | Build(mesh, angular) | Build(angular, XS)")

    def _physics(self):
        L = self.angular.L
        s = self.XS.s if self.XS.s.size else np.zeros((L + 1, self.XS.t.size))
        self._op.set_physics(self.XS.t, np.ascontiguousarray(s.T))
        self.PHYSREADY = True

    def _ready(self, who):
        if not (self.SPACEREADY and self.ANGLEREADY and self.PHYSREADY):
            raise _lib.LionboltError(-3, f"{self.name}%{who} was called, but the operator was not ready.")

    def Destroy(self):
        if self._op is not None:
            self._op.close()
        self._op = None
        self.SPACEREADY = self.ANGLEREADY = self.PHYSREADY = False


class TransportOp(_LinearOperator):
    name = "Transport Operator"

    def MatInv(self, s):
        self._ready("MatInv")
        s.v[:] = self._op.transport_matinv(s.v)

    def MatVec(self, x):
        raise _lib.LionboltError(-1, "TransportOp%MatVec was called, but the operator has yet to be implemented.")


class ScatteringOp(_LinearOperator):
    name = "Scattering Operator"

    def MatVec(self, x):
        self._ready("MatVec")
        x.v[:] = self._op.scatter_matvec(x.v)

    def MatInv(self, s):
        raise _lib.LionboltError(-1, "ScatteringOp%MatInv was called, but this operation is undefined.")


class BoltzmannOp(_LinearOperator):
    """``BoltzmannOp`` with ``solver`` 'GMRES' or 'SI' (the angular set's solver in the reference)."""

    name = "Boltzmann Operator"

    def __init__(self, solver="GMRES", space="psi", restart=10):
        super().__init__()
        self.solver = solver
        self.space = space
        self.restart = restart
        self.stats = None

    def MatVec(self, x):
        self._ready("MatVec")
        x.v[:] = self._op.boltzmann_matvec(x.v)

    def MatInv(self, s):
        self._ready("MatInv")
        out, st = self._op.boltzmann_matinv(s.v, kind=self.solver.lower(), space=self.space, m=self.restart)
        s.v[:] = out
        self.stats = st
