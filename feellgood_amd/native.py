"""ctypes declarations of the engine's two C-ABI libraries.

* ``libfgmi.so``      -- the gfx950 HIP path (include/fgmi.h)
* ``libfgmi_host.so`` -- mesh input, preprocessing, generators, .sol files (include/fgmi_host.h)

Both are built in-tree under ``feellgood_amd/_lib`` by ``feellgood_amd.build``. Loading the HIP
library fails loudly when it is missing: there is no CPU fallback for the hot path.
"""
import ctypes as C
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
HIP_LIB = os.path.join(LIB_DIR, "libfgmi.so")
HOST_LIB = os.path.join(LIB_DIR, "libfgmi_host.so")

FG_OK = 0
FG_SOLVE_FAILED = 1
FG_STEP_REJECTED = 2
FG_ERR_ARG, FG_ERR_DEVICE, FG_ERR_MESH, FG_ERR_STATE, FG_ERR_RANGE = -1, -2, -3, -4, -5
FG_ITER_UNDEFINED, FG_ITER_CONVERGED, FG_ITER_OVERFLOW, FG_ITER_CANNOT_CONVERGE = -1, 0, 1, 2
FG_CURRENT, FG_NEXT, FG_BOTH = 0, 1, 2
FG_DEMAG_DIRECT, FG_DEMAG_FMM = 0, 1
FG_PREC_FP64, FG_PREC_MIXED = 0, 1
FG_KRYLOV_FP64, FG_KRYLOV_MIXED = 0, 1
FG_BATCH_UNION, FG_BATCH_PERSISTENT = 0, 1
ITER_STATUS = {-1: "UNDEFINED", 0: "CONVERGED", 1: "ITER_OVERFLOW", 2: "CANNOT_CONVERGE"}

MU0 = 1.25663706127e-6
GAMMA0 = 1.76085962784e11 * MU0


class VolumeRegion(C.Structure):
    _fields_ = [("alpha", C.c_double), ("Ae", C.c_double), ("Ms", C.c_double), ("K", C.c_double),
                ("uk", C.c_double * 3), ("K3", C.c_double), ("ex", C.c_double * 3),
                ("ey", C.c_double * 3), ("ez", C.c_double * 3)]


class SurfaceRegion(C.Structure):
    _fields_ = [("Ks", C.c_double), ("uk", C.c_double * 3), ("suppress_charges", C.c_int32),
                ("reserved", C.c_int32)]


class Mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_tets", C.c_int32), ("n_tris", C.c_int32),
                ("n_vol_regions", C.c_int32), ("n_surf_regions", C.c_int32), ("reserved", C.c_int32),
                ("pos", C.POINTER(C.c_double)), ("tet_nodes", C.POINTER(C.c_int32)),
                ("tet_region", C.POINTER(C.c_int32)), ("tri_nodes", C.POINTER(C.c_int32)),
                ("tri_region", C.POINTER(C.c_int32)), ("tri_dMs", C.POINTER(C.c_double)),
                ("vol", C.POINTER(VolumeRegion)), ("surf", C.POINTER(SurfaceRegion))]


class SpinVolume(C.Structure):
    _fields_ = [("sigma", C.c_double), ("N0", C.c_double), ("P", C.c_double), ("lsd", C.c_double),
                ("lsf", C.c_double)]


class SpinSurface(C.Structure):
    _fields_ = [("jn", C.c_double), ("uP", C.c_double * 3), ("s", C.c_double * 3)]


class Options(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int32), ("demag_method", C.c_int32),
                ("fmm_order", C.c_int32), ("fmm_leaf", C.c_int32), ("precision", C.c_int32),
                ("device", C.c_int32), ("profile", C.c_int32), ("fmm_far_order", C.c_int32),
                ("krylov", C.c_int32), ("krylov_inner_tol", C.c_double), ("batch_solver", C.c_int32),
                ("batch_block", C.c_int32)]


class SolveInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("residual", C.c_double),
                ("rhs_norm", C.c_double), ("v_max", C.c_double), ("refinements", C.c_int32),
                ("inner_iterations", C.c_int32)]


_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_f32 = C.POINTER(C.c_float)
FG_GUESS_NPARAM = 5959
FG_GUESS_NSCALE = 4        # the scales s_h, s_v, s_out, s_d at the end of the parameter block
FG_GUESS_FEATURE_ROW = 16  # fg_guess_features: u, h, v, d (12 features), the base (3), one pad
# offsets in the parameter block: W1[64][24], b1[64], W2[64][64], b2[64], W3[3][64], b3[3], scales
GUESS_OFFSETS = dict(W1=0, b1=1536, W2=1600, b2=5696, W3=5760, b3=5952, scales=5955)

# name -> (restype, argtypes); every symbol of include/fgmi.h
HIP_SYMBOLS = {
    "fg_default_options": (None, [C.POINTER(Options)]),
    "fg_create": (C.c_int, [C.POINTER(Mesh), C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "fg_create_batch": (C.c_int, [C.POINTER(Mesh), C.c_int, C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "fg_batch_layout": (C.c_int, [C.c_void_p, _i64, _i64, C.POINTER(C.c_int)]),
    "fg_destroy": (C.c_int, [C.c_void_p]),
    "fg_set_state": (C.c_int, [C.c_void_p, C.c_int, _d, _d, _d, _d]),
    "fg_get_state": (C.c_int, [C.c_void_p, C.c_int, _d, _d, _d, _d]),
    "fg_base_projection": (C.c_int, [C.c_void_p, C.c_double]),
    "fg_prepare_elements": (C.c_int, [C.c_void_p, _d, C.c_double, C.c_double]),
    "fg_solve": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(SolveInfo)]),
    "fg_calc_demag": (C.c_int, [C.c_void_p]),
    "fg_energy": (C.c_int, [C.c_void_p, _d, _d]),
    "fg_set_space_field": (C.c_int, [C.c_void_p, _d]),
    "fg_set_spin_accumulation": (C.c_int, [C.c_void_p, _d, C.POINTER(SpinVolume)]),
    "fg_prepare_elements_amp": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "fg_energy_amp": (C.c_int, [C.c_void_p, C.c_double, _d]),
    "fg_gauss_points": (C.c_int, [C.c_void_p, _d]),
    "fg_spin_accumulation": (C.c_int, [C.c_void_p, C.POINTER(SpinVolume), C.POINTER(SpinSurface), _d, _d, _d,
                                       C.c_double, C.c_int32, C.c_int32, _d, C.POINTER(SolveInfo)]),
    "fg_spin_create": (C.c_int, [C.c_void_p, C.POINTER(SpinVolume), C.POINTER(SpinSurface), _d, C.c_double,
                                 C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "fg_spin_solve": (C.c_int, [C.c_void_p, _d, _d, C.c_int32, _d, C.POINTER(SolveInfo)]),
    "fg_spin_solve_ctx": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, _d, C.POINTER(SolveInfo)]),
    "fg_spin_destroy": (C.c_int, [C.c_void_p]),
    "fg_electrostatic_potential": (C.c_int, [C.c_void_p, _d, _d, _d, C.c_double, C.c_int32, C.c_int32, _d,
                                             C.POINTER(SolveInfo)]),
    "fg_electrostatic_potential_noalloc": (C.c_int, [C.c_void_p, _d, _d, _d, C.c_double, C.c_int32, C.c_int32, _d,
                                                     C.POINTER(SolveInfo)]),
    "fg_evolution": (C.c_int, [C.c_void_p]),
    "fg_max_angle": (C.c_int, [C.c_void_p, _d]),
    "fg_step": (C.c_int, [C.c_void_p, _d, C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(SolveInfo), _d,
                          _d]),
    "fg_average": (C.c_int, [C.c_void_p, C.c_int, _d]),
    "fg_get_basis": (C.c_int, [C.c_void_p, _d, _d]),
    "fg_system_size": (C.c_int, [C.c_void_p, _i64, _i64]),
    "fg_get_system": (C.c_int, [C.c_void_p, _i64, _i32, _d, _d]),
    "fg_get_solution": (C.c_int, [C.c_void_p, _d]),
    "fg_kernel_stats": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int, _i64, _d]),
    "fg_reset_kernel_stats": (C.c_int, [C.c_void_p]),
    "fg_fmm_info": (C.c_int, [C.c_void_p, _i64]),
    "fg_guess_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "fg_guess_set": (C.c_int, [C.c_void_p, _f32, C.c_int64]),
    "fg_guess_get": (C.c_int, [C.c_void_p, _f32, C.c_int64]),
    "fg_guess_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "fg_guess_train": (C.c_int, [C.c_void_p, C.c_float, _d, _d]),
    "fg_guess_gradient": (C.c_int, [C.c_void_p, _f32, _d, _d]),
    "fg_guess_features": (C.c_int, [C.c_void_p, _f32]),
    "fg_synchronize": (C.c_int, [C.c_void_p]),
    "fg_fmm_reference": (C.c_int, [_d, C.c_int64, _d, C.c_int64, _d, _d, C.c_int, C.c_int, C.c_int, _d,
                                   _d, C.POINTER(C.c_int)]),
    "fg_last_error": (C.c_char_p, [C.c_void_p]),
}

HOST_SYMBOLS = {
    "fgh_read_msh": (C.c_void_p, [C.c_char_p, C.c_char_p, C.c_int]),
    "fgh_gen_cuboid": (C.c_void_p, [C.c_double] * 6 + [C.c_int] * 3 + [C.c_double, C.c_uint64]),
    "fgh_gen_ellipsoid": (C.c_void_p, [C.c_double] * 7),
    "fgh_gen_cylinder": (C.c_void_p, [C.c_double] * 6),
    "fgh_append": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fgh_free": (None, [C.c_void_p]),
    "fgh_raw_counts": (C.c_int, [C.c_void_p, _i64, _i64, _i64]),
    "fgh_write_msh22": (C.c_int, [C.c_void_p, C.c_char_p]),
    "fgh_prepare": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_char_p),
                              C.POINTER(VolumeRegion), C.c_int, C.POINTER(C.c_char_p),
                              C.POINTER(SurfaceRegion)]),
    "fgh_view": (C.c_int, [C.c_void_p, C.POINTER(Mesh)]),
    "fgh_node_index": (_i32, [C.c_void_p]),
    "fgh_bbox": (C.c_int, [C.c_void_p, _d, _d]),
    "fgh_last_error": (C.c_char_p, []),
    "fgh_eval_expression": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int64, _d, _d]),
    "fgh_read_sol": (C.c_int, [C.c_void_p, C.c_char_p, _d, _d, _d]),
    "fgh_write_sol": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, _d, _d]),
    "fgh_write_sol_s": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, _d, _d, _d]),
}

_libs = {}


def _bind(path, symbols):
    lib = C.CDLL(path)
    for name, (res, args) in symbols.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def host_lib():
    if "host" not in _libs:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built: run feellgood_amd.build.build_all()")
        _libs["host"] = _bind(HOST_LIB, HOST_SYMBOLS)
    return _libs["host"]


def hip_lib():
    """The gfx950 library. Raises when it is not built: the hot path has no CPU fallback."""
    if "hip" not in _libs:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError(f"{HIP_LIB} is not built: run feellgood_amd.build.build_all()")
        _libs["hip"] = _bind(HIP_LIB, HIP_SYMBOLS)
    return _libs["hip"]


def dptr(a):
    """Pointer to a contiguous float64 array (None passes NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_d)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i32)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f32)


def lptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_i64)
