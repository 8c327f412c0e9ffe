"""ctypes binding of libomx_fdtd_hip.so (C-ABI declared in include/omx_fdtd.h).

The library is the HIP engine; there is no fallback.  Loading fails loudly
if the shared object has not been built.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libomx_fdtd_hip.so")

MAX_POLES = 8
MAX_MEDIA = 15


class OmxError(RuntimeError):
    """A non-zero status returned by the engine, with its message."""

    def __init__(self, func, code, msg):
        super().__init__(f"{func} failed with status {code}: {msg}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("npml", C.c_int32),
                ("boundaries", C.c_int32 * 3), ("precision", C.c_int32),
                ("dt", C.c_double), ("dr", C.c_double)]


class LowDesc(C.Structure):
    _fields_ = [("dims", C.c_int32), ("mode", C.c_int32), ("n1", C.c_int32), ("n2", C.c_int32),
                ("npml", C.c_int32), ("boundaries", C.c_int32 * 2), ("precision", C.c_int32),
                ("dt", C.c_double), ("dr", C.c_double)]


class Medium(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_poles", C.c_int32), ("eps_r", C.c_double),
                ("omega_d", C.c_double), ("gamma_d", C.c_double),
                ("omega_k", C.c_double * MAX_POLES), ("gamma_k", C.c_double * MAX_POLES),
                ("a_k", C.c_double * MAX_POLES)]


class MediumCoeffs(C.Structure):
    _fields_ = [("A1", C.c_double), ("A2", C.c_double), ("C1", C.c_double), ("C2", C.c_double),
                ("C3", C.c_double), ("C4", C.c_double), ("inv_eps", C.c_double),
                ("alpha_k", C.c_double * MAX_POLES), ("beta_k", C.c_double * MAX_POLES),
                ("gamma_k", C.c_double * MAX_POLES), ("B1_k", C.c_double * MAX_POLES),
                ("B2_k", C.c_double * MAX_POLES)]


class Shape(C.Structure):
    _fields_ = [("type", C.c_int32), ("medium", C.c_int32), ("p", C.c_double * 15)]


SHAPE_SPHERE, SHAPE_ELLIPSOID, SHAPE_BOX, SHAPE_CYLINDER_Z = 1, 2, 3, 4

_P = C.c_void_p
_I = C.c_int32
_D = C.c_double
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)

_SIGS = {
    "omx_version": ([], C.c_int),
    "omx_medium_library": ([C.c_char_p, _D, C.POINTER(Medium)], C.c_int),
    "omx3d_create": ([C.POINTER(Desc), C.POINTER(_P)], C.c_int),
    "omx3d_destroy": ([_P], C.c_int),
    "omx3d_set_tiling": ([_P, _I, _I, _I], C.c_int),
    "omx3d_get_tiling": ([_P, _IP, _IP, _IP], C.c_int),
    "omx3d_set_addressing": ([_P, _I], C.c_int),
    "omx3d_get_cpml": ([_P, _DP, _DP, _DP, _DP, _DP], C.c_int),
    "omx3d_set_media": ([_P, _I, C.POINTER(Medium), C.POINTER(C.c_int8)], C.c_int),
    "omx3d_set_media_shapes": ([_P, _I, C.POINTER(Medium), _I, C.POINTER(Shape)], C.c_int),
    "omx3d_get_media_map": ([_P, _IP], C.c_int),
    "omx3d_get_medium_coeffs": ([_P, _I, C.POINTER(MediumCoeffs)], C.c_int),
    "omx3d_step_h": ([_P], C.c_int),
    "omx3d_step_e": ([_P, _D, _D], C.c_int),
    "omx3d_add_source": ([_P, _I, _D, _D, _D, _D, _D, _D, _D, _I, _IP, _DP], C.c_int),
    "omx3d_inject_sources": ([_P, _D], C.c_int),
    "omx3d_add_detector": ([_P, _I, _I, _I, _IP], C.c_int),
    "omx3d_set_record_capacity": ([_P, _I], C.c_int),
    "omx3d_record": ([_P], C.c_int),
    "omx3d_read_records": ([_P, _DP, _IP], C.c_int),
    "omx3d_sample": ([_P, _DP], C.c_int),
    "omx3d_set_q_sites": ([_P, _I, _IP, _D, _D], C.c_int),
    "omx3d_set_q_groups": ([_P, _I, _IP, _I, _IP, _IP, _D, _D], C.c_int),
    "omx3d_get_tfsf_line": ([_P, _DP, _DP, _IP, _IP], C.c_int),
    "omx3d_gather_e": ([_P, _DP], C.c_int),
    "omx3d_scatter_j": ([_P, _DP], C.c_int),
    "omx3d_set_two_level": ([_P, _DP, _DP, _DP, _DP], C.c_int),
    "omx3d_two_level_step": ([_P], C.c_int),
    "omx3d_get_two_level": ([_P, _DP, _DP], C.c_int),
    "omx3d_copy_field": ([_P, _I, _DP, _I], C.c_int),
    "omx3d_energy": ([_P, _DP, _DP], C.c_int),
    "omx3d_energy_noalloc": ([_P, _P], C.c_int),
    "omx3d_set_tfsf": ([_P, _IP, _IP, _D, _D, _D, _D, _I], C.c_int),
    "omx3d_run": ([_P, _I, _I, _I], C.c_int),
    "omx3d_set_profiling": ([_P, _I], C.c_int),
    "omx3d_get_profile": ([_P, _DP, _DP, _IP], C.c_int),
    "omx3d_set_source_patch": ([_P, _I], C.c_int),
    "omx3d_get_source_patch": ([_P, _IP, _IP, C.POINTER(C.c_int64)], C.c_int),
    "omx3d_synchronize": ([_P], C.c_int),
    "omx3d_device_bytes": ([_P, C.POINTER(C.c_int64)], C.c_int),
    "omx3d_autotune": ([_P, _I, _IP, _IP, _IP, _DP], C.c_int),
    "omx3d_autotune_step": ([_P, _I, _IP, _DP, _DP], C.c_int),
    "omx3d_set_step_mode": ([_P, _I, _I, _I], C.c_int),
    "omx3d_set_fused_waves": ([_P, _I], C.c_int),
    "omx3d_get_fused_shape": ([_P, _IP, _IP, _IP], C.c_int),
    "omx3d_get_step_mode": ([_P, _IP], C.c_int),
    "omxb_create": ([_I, C.POINTER(_P), C.POINTER(_P)], C.c_int),
    "omxb_run": ([_P, _I, _I, _I, _I], C.c_int),
    "omxb_destroy": ([_P], C.c_int),
    "omxl_create": ([C.c_void_p, C.POINTER(_P)], C.c_int),
    "omxl_destroy": ([_P], C.c_int),
    "omxl_set_media": ([_P, _I, C.c_void_p, C.POINTER(C.c_int8)], C.c_int),
    "omxl_step_h": ([_P], C.c_int),
    "omxl_step_e": ([_P], C.c_int),
    "omxl_add_source": ([_P, _I, _D, _D, _D, _D, _D, _D, _D, _I, _IP, _DP], C.c_int),
    "omxl_inject_sources": ([_P, _D], C.c_int),
    "omxl_add_detector": ([_P, _I, _I, _IP], C.c_int),
    "omxl_sample": ([_P, _DP], C.c_int),
    "omxl_run": ([_P, _I, _I, _I, _DP], C.c_int),
    "omxl_copy_field": ([_P, _I, _DP, _I], C.c_int),
    "omxl_synchronize": ([_P], C.c_int),
    "omxl_device_bytes": ([_P, C.POINTER(C.c_int64)], C.c_int),
}

# every entry point include/omx_fdtd.h declares
EXPORTED = sorted(list(_SIGS) + ["omx_last_error"])

_lib = None


def load_library(path=None):
    """Load the engine (once).  Raises if the shared object is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"libomx_fdtd_hip.so not built at {p}: run __graft_entry__.build()")
    lib = C.CDLL(p)
    lib.omx_last_error.restype = C.c_char_p
    lib.omx_last_error.argtypes = []
    for name, (args, res) in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    if path is None:
        _lib = lib
    return lib


def check(func, rc):
    if rc != 0:
        msg = load_library().omx_last_error().decode(errors="replace")
        raise OmxError(func, rc, msg)


def call(name, *args):
    lib = load_library()
    check(name, getattr(lib, name)(*args))


def dptr(a):
    return a.ctypes.data_as(_DP)


def iptr(a):
    return a.ctypes.data_as(_IP)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
