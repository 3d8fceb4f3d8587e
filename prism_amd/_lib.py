"""ctypes bindings of libprism_amd.so (declared in include/prism_amd.h).

The library is loaded from the package's _build directory; importing a
function from here on a machine where it was not built raises instead of
falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_build", "libprism_amd.so")

PA_OK = 0
PA_ERR_ARG = -1
PA_ERR_HIP = -2
PA_ERR_WORKSPACE = -3
PA_ERR_STREAM = -4
PA_ERR_CAPACITY = -5
PA_F32 = 4
PA_F64 = 8
PA_MODE_NB = 0
PA_MODE_SM = 1

_ERRORS = {PA_ERR_ARG: "invalid argument", PA_ERR_HIP: "HIP runtime error",
           PA_ERR_WORKSPACE: "workspace too small", PA_ERR_STREAM: "malformed stream",
           PA_ERR_CAPACITY: "output buffer too small"}


class PrismError(RuntimeError):
    pass


class Layout(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("n", C.c_int64), ("n4", C.c_int64),
        ("lx", C.c_int64 * 5), ("lxy", C.c_int64 * 5),
        ("cnt", C.c_int64 * 4), ("prefix", C.c_int64 * 5),
        ("dev_align", C.c_int64), ("real_end", C.c_int64), ("code_len", C.c_int64),
        ("dev_len", C.c_int64 * 4), ("dev_stride", C.c_int64 * 4), ("dev_base", C.c_int64 * 4),
        ("dev_size", C.c_int64),
        ("host_stride", C.c_int64 * 4), ("host_base", C.c_int64 * 4), ("host_size", C.c_int64),
        ("chunks_per_plane", C.c_int64 * 4), ("nchunks", C.c_int64), ("trailer_fixed", C.c_int64),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else int(v)
        return out


class BatchFieldDesc(C.Structure):
    """pa_batch_field"""
    _fields_ = [
        ("data", C.c_void_p), ("dtype", C.c_int32), ("mode", C.c_int32), ("rel", C.c_int32),
        ("reserved", C.c_int32), ("eb", C.c_double), ("nx", C.c_uint32), ("ny", C.c_uint32),
        ("nz", C.c_uint32), ("reserved2", C.c_uint32), ("out", C.c_void_p),
    ]


class RetrieveJobDesc(C.Structure):
    """pa_retrieve_job"""
    _fields_ = [
        ("stream", C.c_void_p), ("stream_len", C.c_int64), ("dtype", C.c_int32), ("nx", C.c_uint32),
        ("ny", C.c_uint32), ("nz", C.c_uint32), ("eb", C.c_double), ("rel_eb", C.c_double), ("target", C.c_double),
        ("state", C.c_void_p), ("bp", C.c_void_p), ("qc", C.c_void_p), ("dws", C.c_void_p),
        ("bytes_read", C.c_void_p), ("round", C.c_int32), ("old", C.c_void_p), ("out", C.c_void_p),
    ]


P = C.c_void_p
I32 = C.c_int
I64 = C.c_int64
U32 = C.c_uint32
D = C.c_double
SZ = C.c_size_t
LP = C.POINTER(Layout)

# name -> (restype, argtypes); mirrors include/prism_amd.h
SIGNATURES = {
    "pa_layout_init": (I32, [LP, U32, U32, U32]),
    "pa_compress_bound": (I64, [LP, I32]),
    "pa_extrema": (I32, [P, I32, I64, P, P]),
    "pa_autotune_workspace": (SZ, [LP]),
    "pa_autotune": (I32, [P, I32, LP, P, P, P, SZ, P]),
    "pa_extrema_autotune": (I32, [P, I32, LP, P, D, P, P, P, P, SZ, P]),
    "pa_spline_construct": (I32, [P, I32, LP, D, D, P, P, P, P]),
    "pa_level_bounds": (I32, [P, D, I32, P, P]),
    "pa_spline_construct_dev": (I32, [P, I32, LP, P, P, P, P, P]),
    "pa_spline_reconstruct": (I32, [P, P, I32, LP, D, D, P, P, P, P]),
    "pa_spline_delta_reconstruct": (I32, [P, I32, LP, D, D, P, P, P, P, P]),
    "pa_bitshuffle": (I32, [P, LP, I32, P, P]),
    "pa_inverse_bitshuffle": (I32, [P, LP, I32, P, P]),
    "pa_inverse_bitshuffle_window": (I32, [P, LP, I32, P, P, P, P]),
    "pa_inverse_bitshuffle_window_narrow": (I32, [P, LP, P, P, P, P, P]),
    "pa_spline_reconstruct_narrow": (I32, [P, P, P, I32, LP, D, D, P, P, P, P]),
    "pa_spline_delta_reconstruct_narrow": (I32, [P, P, I32, LP, D, D, P, P, P, P, P]),
    "pa_inverse_bitshuffle_window_deferred": (I32, [P, LP, P, P, P, P, P]),
    "pa_spline_reconstruct_planes": (I32, [P, P, P, P, P, P, P, I32, LP, D, D, P, P, P, P]),
    "pa_encode_workspace": (SZ, [LP]),
    "pa_lossless_encode": (I32, [P, LP, P, P, P, P, SZ, P]),
    "pa_compress_merge": (I32, [P, LP, I32, P, P, P, P, P, P, P]),
    "pa_decode_workspace": (SZ, [LP]),
    "pa_lossless_decode": (I32, [P, I64, LP, P, P, SZ, P]),
    "pa_lossless_decode_window": (I32, [P, I64, LP, P, P, P, P, I32, P, SZ, P]),
    "pa_noalloc_scratch_bytes": (SZ, []),
    "pa_lossless_decode_noalloc": (I32, [P, I64, LP, P, P, P]),
    "pa_strategy_workspace": (SZ, []),
    "pa_find_strategy": (I32, [P, P, P, P, D, D, P, SZ, P]),
    "pa_quality_workspace": (SZ, [I64]),
    "pa_quality": (I32, [P, P, I32, I64, P, P, SZ, P]),
    "pa_batch_workspace": (SZ, [P, I32]),
    "pa_batch_plan_create": (I32, [P, I32, P, SZ, P, I32, P, C.POINTER(C.c_void_p)]),
    "pa_batch_plan_rebind": (I32, [P, P, I32, P]),
    "pa_batch_run": (I32, [P, P]),
    "pa_batch_plan_launches": (I32, [P]),
    "pa_batch_plan_autotune": (I32, [P, I32, P]),
    "pa_batch_plan_params": (I32, [P, P, P]),
    "pa_batch_plan_destroy": (I32, [P]),
    "pa_retrieve_table_bytes": (SZ, [I32]),
    "pa_retrieve_workspace": (SZ, [I32]),
    "pa_retrieve_batch": (I32, [P, I32, P, SZ, P, SZ, P]),
    "pa_version": (C.c_char_p, []),
    "pa_device_count": (I32, []),
}

_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PrismError(f"{LIB_PATH} is missing: build it with `python -m prism_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != PA_OK:
        raise PrismError(f"{what}: {_ERRORS.get(rc, rc)}")


def layout(nx: int, ny: int = 1, nz: int = 1) -> Layout:
    L = Layout()
    check(lib().pa_layout_init(C.byref(L), nx, ny, nz), "pa_layout_init")
    return L
