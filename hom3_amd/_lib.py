"""ctypes binding of libhom3g.so (the C ABI in include/hom3g.h).

The shared library is built in-tree by ``hom3_amd.build``; importing this
module never falls back to anything else: if the library is missing or no
HIP device is present, the calls fail with an explicit error.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhom3g.so")

OK = 0
INVALID_ID = (1 << 63) - 1

EULER, CNS, HEAT = 0, 1, 2
FLUX_AUSM, FLUX_LLF, FLUX_THREE_POINT = 0, 1, 2
TI_EULER_FWD, TI_HEUN_RK2, TI_SSP_RK3 = 0, 1, 2
RECON_NONE, RECON_MUSCL = 0, 1
LIM_ZERO, LIM_MINMOD, LIM_VANLEER = 0, 1, 2
FP64, FP32_STATE = 0, 1
KERNEL_AUTO, KERNEL_GENERIC, KERNEL_BRICK = 0, 1, 2
GEOM_NONE, GEOM_EDGE, GEOM_SQUARE, GEOM_SPHERE = 0, 1, 2, 3
(BC_NEUMANN, BC_DIRICHLET, BC_ISENTROPIC_VORTEX, BC_INFLOW, BC_OUTFLOW,
 BC_ADIABATIC_NOSLIP, BC_ADIABATIC_SLIP, BC_ISOTHERMAL_NOSLIP, BC_HEAT_DIRICHLET,
 BC_HEAT_NEUMANN, BC_PERIODIC, BC_CHT_FLUID, BC_CHT_SOLID, BC_EULER_COUPLING,
 BC_HOST_CALLBACK) = range(15)
SLOPE_DEFAULT, SLOPE_WALL = 0, 1
ERR_DIVERGED = -4
(IC_CONSTANT_PV, IC_SHOCK_TUBE, IC_ISENTROPIC_VORTEX, IC_SPHERE, IC_TGV,
 IC_FOUR_QUADRANT, IC_HEAT_CONSTANT, IC_HEAT_SINE, IC_LINEAR_PV) = range(9)


# host callbacks of the C ABI (hom3g_bc_callback, hom3g_ic_callback)
BC_CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, C.c_int64, C.POINTER(C.c_int64),
                          C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.POINTER(C.c_double), C.POINTER(C.c_double))
IC_CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


class Geometry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("invert", C.c_int32),
                ("a", C.c_double * 3), ("b", C.c_double * 3)]


class Physics(C.Structure):
    _fields_ = [("physics", C.c_int32), ("nd", C.c_int32), ("flux", C.c_int32),
                ("ti", C.c_int32), ("recon", C.c_int32), ("limiter", C.c_int32),
                ("gamma", C.c_double), ("cfl", C.c_double), ("cfl_viscous", C.c_double),
                ("Rspecific", C.c_double), ("T0", C.c_double), ("rho0", C.c_double),
                ("Minfty", C.c_double), ("ReInfty", C.c_double), ("Pr0", C.c_double),
                ("Sutherland0", C.c_double), ("diffusivity", C.c_double),
                ("T_ref", C.c_double), ("time_end", C.c_double), ("max_cells", C.c_int64),
                ("precision", C.c_int32), ("kernel", C.c_int32)]


# name -> (restype, argtypes); every symbol declared in include/hom3g.h
_vp, _i32, _i64, _dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
_P = C.POINTER
SIGNATURES = {
    "hom3g_last_error": (C.c_char_p, []),
    "hom3g_version": (C.c_int, []),
    "hom3g_create": (C.c_int, [_P(_vp), _i32]),
    "hom3g_destroy": (C.c_int, [_vp]),
    "hom3g_synchronize": (C.c_int, [_vp]),
    "hom3g_grid_create": (C.c_int, [_vp, _i32, _P(_dbl), _P(_dbl), _i32, _i32, _P(_i32)]),
    "hom3g_grid_set_tree": (C.c_int, [_vp, _i32, _i64, _P(_i64), _P(_i64)]),
    "hom3g_grid_refine_region": (C.c_int, [_vp, _i32, _i32, _P(Geometry), C.c_double, _P(_i64)]),
    "hom3g_grid_balance": (C.c_int, [_vp, _i32, _P(_i64)]),
    "hom3g_grid_size": (C.c_int, [_vp, _i32, _P(_i64)]),
    "hom3g_grid_tree": (C.c_int, [_vp, _i32, _P(_i64), _P(_i64)]),
    "hom3g_grid_neighbors": (C.c_int, [_vp, _i32, _P(_i64)]),
    "hom3g_grid_geometry": (C.c_int, [_vp, _i32, _P(_dbl), _P(_dbl)]),
    "hom3g_solver_create": (C.c_int, [_vp, _i32, _i32, _P(Physics), _P(_i32)]),
    "hom3g_solver_set_domain": (C.c_int, [_vp, _i32, _P(Geometry), _i32]),
    "hom3g_solver_set_domain_mask": (C.c_int, [_vp, _i32, _vp, _i64]),
    "hom3g_solver_append_bc": (C.c_int, [_vp, _i32, _P(Geometry), _i32, _P(_dbl), _i32, _i32, _dbl]),
    "hom3g_solver_build": (C.c_int, [_vp, _i32]),
    "hom3g_solver_impose_ic": (C.c_int, [_vp, _i32, _i32, _P(_dbl), _i32]),
    "hom3g_solver_set_perturbation": (C.c_int, [_vp, _i32, C.c_uint64, C.c_double]),
    "hom3g_solver_fine_faces": (C.c_int, [_vp, _i32, _vp, _P(_i64)]),
    "hom3g_solver_counts": (C.c_int, [_vp, _i32, _P(_i64), _P(_i64)]),
    "hom3g_solver_info": (C.c_int, [_vp, _i32, _P(_i32), _P(_i32), _P(_i32)]),
    "hom3g_solver_kernel": (C.c_int, [_vp, _i32, _P(_i32)]),
    "hom3g_solver_topology": (C.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "hom3g_solver_upload_state": (C.c_int, [_vp, _i32, _P(_dbl), _i64]),
    "hom3g_solver_download_state": (C.c_int, [_vp, _i32, _P(_dbl), _i64]),
    "hom3g_solver_reset_time": (C.c_int, [_vp, _i32, _dbl, _dbl, _i64]),
    "hom3g_solver_set_source": (C.c_int, [_vp, _i32, _P(_dbl), _i64]),
    "hom3g_solver_solve": (C.c_int, [_vp, _i32, _i32]),
    "hom3g_solver_min_dt": (C.c_int, [_vp, _i32, _P(_dbl)]),
    "hom3g_solver_force_dt": (C.c_int, [_vp, _i32, _dbl]),
    "hom3g_solver_time": (C.c_int, [_vp, _i32, _P(_dbl), _P(_dbl), _P(_i64)]),
    "hom3g_solver_conservation_sums": (C.c_int, [_vp, _i32, _P(_dbl)]),
    "hom3g_solver_append_bc_host": (C.c_int, [_vp, _i32, _P(Geometry), BC_CALLBACK, _vp, _i32]),
    "hom3g_solver_impose_ic_host": (C.c_int, [_vp, _i32, IC_CALLBACK, _vp]),
    "hom3g_solver_apply_bcs": (C.c_int, [_vp, _i32]),
    "hom3g_solver_conservation_sums_noalloc": (C.c_int, [_vp, _i32, _vp]),
    "hom3g_solver_check_state": (C.c_int, [_vp, _i32, _P(_i64)]),
    "hom3g_solver_time_solve": (C.c_int, [_vp, _i32, _i32, _P(C.c_float)]),
    "hom3g_solver_profile_solve": (C.c_int, [_vp, _i32, _i32, _P(C.c_float), _P(C.c_float), _P(_i32)]),
    "hom3g_solver_destroy": (C.c_int, [_vp, _i32]),
    "hom3g_copy_calibration": (C.c_int, [_vp, _i64, _P(C.c_float)]),
    "hom3g_copy_bandwidth": (C.c_int, [_vp, _i64, C.c_int, C.c_int, C.c_int, C.c_int, _P(C.c_float)]),
    "hom3g_couple_cht_step": (C.c_int, [_vp, _i32, _i32, _dbl, _i32]),
    "hom3g_couple_same_step": (C.c_int, [_vp, _i32, _i32, _i32]),
    "hom3g_cht_objective": (C.c_int, [_vp, _i32, _i32, C.c_double, _i32, _P(C.c_double), _P(C.c_double),
                                      _P(C.c_double)]),
    "hom3g_cht_tangent": (C.c_int, [_vp, _i32, _i32, C.c_double, _i32, _P(C.c_double), _P(C.c_double),
                                    _P(C.c_double), _P(C.c_double)]),
    "hom3g_cht_adjoint": (C.c_int, [_vp, _i32, _i32, C.c_double, _i32, _P(C.c_double), _i32, _P(C.c_double),
                                    _P(C.c_double), _P(C.c_float), _P(C.c_float)]),
    "hom3g_set_checkpoint_memory": (C.c_int, [_vp, _i32]),
    "hom3g_cht_misfit_objective": (C.c_int, [_vp, _i32, _i32, C.c_double, _i32, _P(C.c_double), _P(C.c_double),
                                             _P(C.c_double)]),
    "hom3g_cht_misfit_tangent": (C.c_int, [_vp, _i32, _i32, C.c_double, _i32, _P(C.c_double), _P(C.c_double),
                                           _P(C.c_double), _P(C.c_double), _P(C.c_double)]),
    "hom3g_cht_misfit_adjoint": (C.c_int, [_vp, _i32, _i32, C.c_double, _i32, _P(C.c_double), _i32,
                                           _P(C.c_double), _P(C.c_double), _P(C.c_double), _P(C.c_float),
                                           _P(C.c_float)]),
    "hom3g_batch_create": (C.c_int, [_vp, _P(_i32), _i32, _i32, _P(_i32)]),
    "hom3g_batch_set_config": (C.c_int, [_vp, _i32, _i32, _i32, _i32]),
    "hom3g_batch_get_config": (C.c_int, [_vp, _i32, _P(_i32), _P(_i32), _P(_i32), _P(_i64)]),
    "hom3g_batch_autotune": (C.c_int, [_vp, _i32, _P(_i32), _P(_i32), _P(_i32), _P(C.c_float)]),
    "hom3g_batch_solve": (C.c_int, [_vp, _i32, _i32, _i32]),
    "hom3g_batch_time_solve": (C.c_int, [_vp, _i32, _i32, _i32, _P(C.c_float)]),
    "hom3g_batch_counts": (C.c_int, [_vp, _i32, _P(_i32), _P(_i64)]),
    "hom3g_batch_download_state": (C.c_int, [_vp, _i32, _i32, _P(C.c_double), _i64]),
    "hom3g_batch_time": (C.c_int, [_vp, _i32, _i32, _P(C.c_double), _P(C.c_double), _P(_i64)]),
    "hom3g_batch_destroy": (C.c_int, [_vp, _i32]),
}

_lib = None


class Hom3gError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hom3g error {code}: {msg}")
        self.code = code


def lib():
    """Load libhom3g.so (never a fallback).  Raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m hom3_amd.build`")
        # One HIP runtime per process: torch bundles libamdhip64 under a file
        # name (libamdhip64.so) other than the soname libhom3g.so needs
        # (libamdhip64.so.7).  Loaded first, torch's copy provides that soname
        # to libhom3g; loaded after libhom3g, it would be a second runtime
        # that finds no GPU once the first holds it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise Hom3gError(rc, lib().hom3g_last_error().decode())
    return rc
