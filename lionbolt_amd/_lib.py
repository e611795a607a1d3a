"""ctypes binding of liblbhip.so (the C ABI declared in include/lionbolt.h).

The library is loaded from this package directory only; a missing library is a
hard error, there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbhip.so")

LB_OK = 0
ERRORS = {-1: "LB_E_ARG", -2: "LB_E_STAGNANT", -3: "LB_E_NOT_READY", -4: "LB_E_HIP", -5: "LB_E_NOMEM",
          -6: "LB_E_NAN", -7: "LB_E_MAXIT", -8: "LB_E_BREAKDOWN"}
SOLVER_GMRES, SOLVER_SI, SOLVER_FGMRES = 0, 1, 2
SPACE_PSI, SPACE_MOMENT = 0, 1
SCATTER_MOMENT, SCATTER_DENSE = 0, 1
SWEEP_LEVELS, SWEEP_GRAPH, SWEEP_DIRECTION, SWEEP_AUTO, SWEEP_PERSISTENT, SWEEP_DIRECTION_LDS = 0, 1, 2, 3, 4, 5
FLUX_NATURAL, FLUX_SWEEP_ORDER, FLUX_SWEEP_ORDER_SOURCE, FLUX_SWEEP_ORDER_FP32 = 0, 1, 2, 3

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i32 = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p
_f = ctypes.POINTER(ctypes.c_float)
FMT_FP32, FMT_BF16, FMT_SPLIT_BF16, FMT_FP16, FMT_FP64 = 0, 1, 2, 3, 4
MLP_PARAMS = 10576


class MeshDesc(ctypes.Structure):
    _fields_ = [("NE", ctypes.c_int64), ("NK", ctypes.c_int64), ("rg", _d), ("node", _i64), ("nbr", _i64),
                ("mat", _i64), ("index_base", ctypes.c_int)]


class QuadDesc(ctypes.Structure):
    _fields_ = [("NI", ctypes.c_int64), ("k", _d), ("w", _d), ("L", ctypes.c_int)]


class SolverOpts(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("space", ctypes.c_int), ("restart", ctypes.c_int), ("maxit", ctypes.c_int),
                ("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("verbose", ctypes.c_int)]


CUTOUT_NONE, CUTOUT_CIRCLE, CUTOUT_RECTANGLE = 0, 1, 2


class BeamDesc(ctypes.Structure):
    _fields_ = [("axis", ctypes.c_double * 3), ("origin", ctypes.c_double * 3), ("e1", ctypes.c_double * 3),
                ("e2", ctypes.c_double * 3), ("cutout", ctypes.c_int), ("params", ctypes.c_double * 4),
                ("n", ctypes.c_double)]


class SolveStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterates", ctypes.c_int), ("matvecs", ctypes.c_int),
                ("sweeps", ctypes.c_int), ("final_norm", ctypes.c_double), ("true_residual", ctypes.c_double),
                ("seconds", ctypes.c_double), ("fallback", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "lb_op_create": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(QuadDesc), ctypes.POINTER(_vp)]),
    "lb_op_set_physics": (ctypes.c_int, [_vp, ctypes.c_int, _d, ctypes.c_int, _d]),
    "lb_op_set_physics_groups": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _d, ctypes.c_int, _d]),
    "lb_op_set_scatter_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lb_op_set_sweep_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lb_op_build_delta_down": (ctypes.c_int, [_vp, ctypes.c_int, _d]),
    "lb_op_sweep_choice": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)]),
    "lb_op_set_workgroups_per_cu": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lb_op_set_group_batch": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lb_op_set_problems": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
    "lb_batch_solve_dev": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(SolverOpts), _vp, ctypes.POINTER(SolveStats)]),
    "lb_op_set_preconditioner": (ctypes.c_int, [_vp, _f, ctypes.c_int, ctypes.c_int]),
    "lb_mlp_apply_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "lb_mlp_apply_noalloc_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "lb_op_set_preconditioner_threshold": (ctypes.c_int, [_vp, ctypes.c_double]),
    "lb_op_set_moment_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lb_beam_uncollided_dev": (ctypes.c_int, [_vp, ctypes.POINTER(BeamDesc), _d, _vp]),
    "lb_beam_uncollided_range_dev": (ctypes.c_int, [_vp, ctypes.POINTER(BeamDesc), _d, ctypes.c_int64, ctypes.c_int64,
                                                    _vp]),
    "lb_beam_moments_dev": (ctypes.c_int, [_vp, ctypes.POINTER(BeamDesc), _vp, _vp, _vp]),
    "lb_beam_angular_fluence_dev": (ctypes.c_int, [_vp, ctypes.POINTER(BeamDesc), _vp, _vp]),
    "lb_beam_source": (ctypes.c_int, [_vp, ctypes.POINTER(BeamDesc), _d, ctypes.c_int64]),
    "lb_beam_add_angular_fluence": (ctypes.c_int, [_vp, ctypes.POINTER(BeamDesc), _d, _d, ctypes.c_int64]),
    "lb_mlp_features_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_double)]),
    "lb_mlp_train_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, _f, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_float, ctypes.c_uint64, _f]),
    "lb_mlp_grad_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _f, ctypes.c_int, _f, _f]),
    "lb_moment_rhs_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "lb_moment_matvec_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "lb_discrete_source_dev": (ctypes.c_int, [_vp, _vp, _vp]),
    "lb_gmres_snapshots_dev": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(SolverOpts), _vp, _vp, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(SolveStats)]),
    "lb_op_set_direction_block": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lb_transport_matinv": (ctypes.c_int, [_vp, _d, ctypes.c_int64]),
    "lb_scatter_matvec": (ctypes.c_int, [_vp, _d, ctypes.c_int64]),
    "lb_boltzmann_matvec": (ctypes.c_int, [_vp, _d, ctypes.c_int64]),
    "lb_boltzmann_matinv": (ctypes.c_int, [_vp, _d, ctypes.c_int64, ctypes.POINTER(SolverOpts),
                                           ctypes.POINTER(SolveStats)]),
    "lb_fluence": (ctypes.c_int, [_vp, _d, _d]),
    "lb_op_get_schedule": (ctypes.c_int, [_vp, _i32, _i32]),
    "lb_schedule_host": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(QuadDesc), _i32, _i32, _d]),
    "lb_trace_plan_host": (ctypes.c_int, [ctypes.POINTER(MeshDesc), ctypes.POINTER(QuadDesc), ctypes.c_int,
                                          ctypes.c_int, _i32, _i32, _i32, _i32, _i64, _i32, _i64, _vp]),
    "lb_op_get_normals": (ctypes.c_int, [_vp, _d]),
    "lb_op_info": (ctypes.c_int, [_vp, _i64, _i64, _i64, _i64]),
    "lb_transport_matinv_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int]),
    "lb_moments_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "lb_moments_to_discrete_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "lb_op_set_flux_order": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lb_op_sweep_order_choice": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_int)]),
    "lb_transport_matinv_order_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int]),
    "lb_moments_to_discrete_order_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "lb_moments_order_dev": (ctypes.c_int, [_vp, _vp, _vp]),
    "lb_transport_matinv_order32_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int]),
    "lb_moments_order32_dev": (ctypes.c_int, [_vp, _vp, _vp]),
    "lb_flux_to_natural_dev": (ctypes.c_int, [_vp, _vp, _vp]),
    "lb_multigroup_solve_dev": (ctypes.c_int, [_vp, _vp, _d, ctypes.POINTER(SolverOpts), _vp,
                                               ctypes.POINTER(SolveStats)]),
    "lb_pack_dev": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp]),
    "lb_gemm_d_dev": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _vp, _vp, _vp,
                                     _vp, _vp]),
    "lb_gemm_m_dev": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _vp, _vp, _vp,
                                     _vp, _vp]),
    "lb_op_synchronize": (ctypes.c_int, [_vp]),
    "lb_last_sweep_ms": (ctypes.c_int, [_vp, _d]),
    "lb_op_destroy": (None, [_vp]),
    "lb_last_error": (ctypes.c_char_p, []),
}

_LIB = None


class LionboltError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                              "(make -C lionbolt_amd/csrc). There is no CPU fallback.")
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
        # Loaded first, it is the one the library binds to, so torch tensors and the
        # library share one runtime; loaded after the library it would find the
        # system runtime already in place and fail to initialise.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != LB_OK:
        raise LionboltError(rc, lib().lb_last_error().decode())
    return rc


def dptr(a):
    return a.ctypes.data_as(_d)


def i64ptr(a):
    return a.ctypes.data_as(_i64)


def i32ptr(a):
    return a.ctypes.data_as(_i32)
