"""ctypes binding of libwavi_hip.so (declared in include/wavi_hip.h).

The library is the only compute path of this package: if it is missing, every
solve raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libwavi_hip.so")

WAVI_OK = 0
WAVI_EARG = -1
WAVI_EDIM = -2
WAVI_ESTATE = -3
WAVI_EHIP = 1

# coarse-CG modes (wavi_set_coarse_cg)
WAVI_CG_AUTO = 0
WAVI_CG_LAUNCHES = 1
WAVI_CG_WORKGROUP = 2

PRECISION = {"fp64": 0, "mixed": 1}

REPORT_MAX_PICARD = 64

KT = {"cg_apply": 0, "cg_dir": 1, "cg_update": 2, "colour": 3, "cg_iter_mixed": 4, "cg_iter_fp64": 5}

# enum wavi_field, in header order
FIELD_NAMES = [
    "H_B", "H_H", "H_S", "H_DHDT", "H_ACCUMULATION", "H_BASAL_MELT", "H_HAF",
    "H_GROUNDED_FRACTION", "H_DSDH", "H_SHELF_STRAIN_RATE", "H_AV_SPEED", "H_U",
    "H_V", "H_US", "H_VS", "H_UB", "H_VB", "H_BED_SPEED", "H_WEERTMAN_C",
    "H_BETA", "H_BETA_EFF", "H_TAU_BED", "H_ETA_AV", "H_QUAD_F1", "H_QUAD_F2",
    "P_WEERTMAN_C", "P_ACCUMULATION_RATE",
    "U_S", "U_H", "U_GROUNDED_FRACTION", "U_BETA_EFF", "U_U", "U_RESIDUAL",
    "V_S", "V_H", "V_GROUNDED_FRACTION", "V_BETA_EFF", "V_V", "V_RESIDUAL",
    "S_ETA", "S_GLEN_B",
    "WU_WAVELETS", "WU_MASK", "WU_CORRECTION_COARSE",
    "WV_WAVELETS", "WV_MASK", "WV_CORRECTION_COARSE",
]
FIELD = {name: i for i, name in enumerate(FIELD_NAMES)}
NFIELDS = len(FIELD_NAMES)

_u8p = ctypes.POINTER(ctypes.c_uint8)
_dp = ctypes.POINTER(ctypes.c_double)


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nsigma", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("h_mask", _u8p), ("h_isfixed", _u8p), ("u_iszero", _u8p),
        ("v_iszero", _u8p), ("u_isfixed", _u8p), ("v_isfixed", _u8p),
        ("sigma", _dp), ("quadrature_weights", _dp),
    ]


class PhysParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("g", ctypes.c_double),
        ("density_ice", ctypes.c_double), ("density_ocean", ctypes.c_double),
        ("glen_n", ctypes.c_double), ("glen_reg_strain_rate", ctypes.c_double),
        ("weertman_m", ctypes.c_double), ("weertman_reg_speed", ctypes.c_double),
        ("sea_level_wrt_geoid", ctypes.c_double), ("minimum_thickness", ctypes.c_double),
        ("uniform_melt_rate", ctypes.c_double), ("evolve_shelves", ctypes.c_int32),
    ]


class SolverParamsC(ctypes.Structure):
    _fields_ = [
        ("n_iter_viscosity", ctypes.c_int32), ("tol_picard", ctypes.c_double),
        ("maxiter_picard", ctypes.c_int32), ("tol_coarse", ctypes.c_double),
        ("maxiter_coarse", ctypes.c_int32), ("levels", ctypes.c_int32),
        ("wavelet_threshold", ctypes.c_double), ("nsmooth", ctypes.c_int32),
        ("smoother_omega", ctypes.c_double), ("stencil_margin", ctypes.c_int32),
        ("super_implicitness", ctypes.c_double),
    ]


class SolveReport(ctypes.Structure):
    _fields_ = [
        ("picard_iters", ctypes.c_int32), ("converged", ctypes.c_int32),
        ("rel_resid", ctypes.c_double), ("coarse_cg_total", ctypes.c_int32),
        ("refinements", ctypes.c_int32),
        ("coarse_cg_iters", ctypes.c_int32 * REPORT_MAX_PICARD),
        ("rel_resid_hist", ctypes.c_double * REPORT_MAX_PICARD),
    ]

    def as_dict(self):
        n = min(self.picard_iters, REPORT_MAX_PICARD)
        return {
            "picard_iters": self.picard_iters,
            "converged": bool(self.converged),
            "rel_resid": self.rel_resid,
            "coarse_cg_total": self.coarse_cg_total,
            "refinements": self.refinements,
            "coarse_cg_iters": list(self.coarse_cg_iters[:n]),
            "rel_resid_hist": list(self.rel_resid_hist[:n]),
        }


class AdjointReport(ctypes.Structure):
    _fields_ = [
        ("n_picard", ctypes.c_int32), ("fwd_vcycles", ctypes.c_int32), ("adj_vcycles", ctypes.c_int32),
        ("unconverged", ctypes.c_int32), ("max_fwd_rel", ctypes.c_double), ("max_adj_rel", ctypes.c_double),
        ("coarse_cg_iters", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# exported symbols and their ctypes signatures (kept in step with the header)
SIGNATURES = {
    "wavi_ctx_create": (ctypes.c_int, [ctypes.POINTER(GridDesc), ctypes.POINTER(PhysParams),
                                       ctypes.POINTER(SolverParamsC), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "wavi_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "wavi_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "wavi_set_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "wavi_field_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "wavi_set_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_size_t]),
    "wavi_get_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_size_t]),
    "wavi_init_quad_f2": (ctypes.c_int, [ctypes.c_void_p]),
    "wavi_state_save": (ctypes.c_int, [ctypes.c_void_p]),
    "wavi_state_restore": (ctypes.c_int, [ctypes.c_void_p]),
    "wavi_update_velocities": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolveReport)]),
    "wavi_misfit_gradient_batch": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                  ctypes.POINTER(_dp), ctypes.POINTER(_dp), _dp,
                                                  ctypes.POINTER(_dp), ctypes.POINTER(AdjointReport),
                                                  ctypes.POINTER(ctypes.c_int)]),
    "wavi_update_state_batch": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(SolveReport), ctypes.POINTER(ctypes.c_int)]),
    "wavi_after_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "wavi_update_velocities_batch": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                                    ctypes.POINTER(SolveReport),
                                                    ctypes.POINTER(ctypes.c_int)]),
    "wavi_update_velocities_lockstep": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                                       ctypes.POINTER(SolveReport), ctypes.POINTER(ctypes.c_int)]),
    "wavi_update_velocities_lockstep_noalloc": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                                               ctypes.POINTER(SolveReport),
                                                               ctypes.POINTER(ctypes.c_int)]),
    "wavi_schwarz_precondition": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_double)]),
    "wavi_schwarz_refresh": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int]),
    "wavi_inner_update": (ctypes.c_int, [ctypes.c_void_p]),
    "wavi_update_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolveReport)]),
    "wavi_update_geometry": (ctypes.c_int, [ctypes.c_void_p]),
    "wavi_update_thickness": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "wavi_run_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(SolveReport)]),
    "wavi_update_wavelets": (ctypes.c_int, [ctypes.c_void_p]),
    "wavi_misfit_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                            ctypes.POINTER(ctypes.c_double), _dp, ctypes.POINTER(AdjointReport)]),
    "wavi_op_apply": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.c_int]),
    "wavi_rhs": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "wavi_op_diag": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "wavi_vcycle": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int32)]),
    "wavi_haar": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "wavi_set_smoother_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "wavi_smoother_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "wavi_set_coarse_cg": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "wavi_set_linear_coarse_jacobi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "wavi_coarse_diag": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "wavi_coarse_cg": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "wavi_profile_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "wavi_profile_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_int64)]),
    "wavi_math_probe": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, ctypes.c_size_t]),
}

_LIB = None


def load(path=LIB_PATH):
    """Load libwavi_hip.so; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(
            "libwavi_hip.so not found at %s: build it first (python -c "
            "'import __graft_entry__ as g; g.build()')" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def u8ptr(a):
    if a is None:
        return None
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p)
