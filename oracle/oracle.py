"""ctypes bindings of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: imported only by the test suite, by the smoke check and
by the CPU baseline timing in bench.py.  The product package never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
CS = 16384


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


def build(force: bool = False) -> str:
    """Compiles the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
I64 = C.c_int64
D = C.c_double
U32 = C.c_uint32
INT = C.c_int


def _declare(L):
    L.po_layout_init.argtypes = [C.POINTER(Layout), U32, U32, U32]
    L.po_lm_index.argtypes = [C.POINTER(Layout), U32, U32, U32]
    L.po_lm_index.restype = I64
    L.po_alpha.argtypes = [D]
    L.po_alpha.restype = D
    for s in ("f32", "f64"):
        getattr(L, f"po_extrema_{s}").argtypes = [P, I64, P]
        getattr(L, f"po_spline_construct_{s}").argtypes = [P, C.POINTER(Layout), D, D, P, P, P, INT]
        getattr(L, f"po_spline_reconstruct_{s}").argtypes = [P, P, C.POINTER(Layout), D, D, P, INT, P, INT]
        getattr(L, f"po_spline_delta_reconstruct_{s}").argtypes = [P, C.POINTER(Layout), D, D, P, INT, P, P, INT]
        getattr(L, f"po_autotune_{s}").argtypes = [P, C.POINTER(Layout), D, P, P]
        getattr(L, f"po_autotune_{s}").restype = INT
        getattr(L, f"po_compress_{s}").argtypes = [P, U32, U32, U32, D, INT, INT, P, I64, P, INT]
        getattr(L, f"po_compress_{s}").restype = I64
        getattr(L, f"po_decompress_{s}").argtypes = [P, I64, U32, U32, U32, D, INT, P, INT]
        getattr(L, f"po_decompress_{s}").restype = INT
        getattr(L, f"po_progressive_round_{s}").argtypes = [P, I64, U32, U32, U32, D, INT, D, INT, INT,
                                                             P, P, P, P, P, INT]
        getattr(L, f"po_progressive_round_{s}").restype = INT
    L.po_pick_flags.argtypes = [P, P]
    L.po_bitshuffle.argtypes = [P, C.POINTER(Layout), INT, P]
    L.po_inverse_bitshuffle.argtypes = [P, C.POINTER(Layout), INT, P]
    L.po_inverse_bitshuffle_window.argtypes = [P, C.POINTER(Layout), INT, P, P, P]
    L.po_rze_encode_chunk.argtypes = [P, INT, P]
    L.po_rze_encode_chunk.restype = INT
    L.po_rze_decode_chunk.argtypes = [P, INT, P]
    L.po_rze_decode_chunk.restype = INT
    L.po_stream_bound.argtypes = [C.POINTER(Layout)]
    L.po_stream_bound.restype = I64
    L.po_lossless_encode.argtypes = [P, C.POINTER(Layout), P, P]
    L.po_lossless_encode.restype = I64
    L.po_lossless_decode.argtypes = [P, I64, C.POINTER(Layout), P]
    L.po_lossless_decode.restype = INT
    L.po_lossless_decode_window.argtypes = [P, I64, C.POINTER(Layout), P, P, INT, P]
    L.po_lossless_decode_window.restype = I64
    L.po_find_strategy.argtypes = [P, P, P, D, D, P]
    L.po_find_strategy.restype = INT
    L.po_strategy_search.argtypes = [P, C.c_uint64, P, P]
    L.po_strategy_cost.argtypes = [INT, INT]
    L.po_strategy_cost.restype = C.c_uint64
    L.po_compress_bound.argtypes = [U32, U32, U32, INT]
    L.po_compress_bound.restype = I64
    L.po_parse_trailer.argtypes = [P, I64, C.POINTER(Layout), INT, P, P, P, P]
    L.po_parse_trailer.restype = INT


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def sfx(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def layout(nx: int, ny: int = 1, nz: int = 1) -> Layout:
    L = Layout()
    lib().po_layout_init(C.byref(L), nx, ny, nz)
    return L


def lm_index(L: Layout, gx: int, gy: int, gz: int) -> int:
    return lib().po_lm_index(C.byref(L), gx, gy, gz)


def alpha(rel_eb: float) -> float:
    return lib().po_alpha(rel_eb)


def extrema(data: np.ndarray):
    out = np.zeros(2, np.float64)
    getattr(lib(), f"po_extrema_{sfx(data.dtype)}")(ptr(data), data.size, ptr(out))
    return float(out[0]), float(out[1])


def nthreads_default() -> int:
    return max(1, min(os.cpu_count() or 1, 16))


def spline_construct(data: np.ndarray, eb: float, rel_eb: float, flags, nthreads=None):
    """data: [nz][ny][nx] array. Returns (qc[code_len] int32, anchors[n4])."""
    nz, ny, nx = data.shape
    L = layout(nx, ny, nz)
    qc = np.zeros(L.code_len, np.int32)
    anchors = np.zeros(L.n4, data.dtype)
    fl = np.asarray(flags, np.uint8)
    getattr(lib(), f"po_spline_construct_{sfx(data.dtype)}")(
        ptr(np.ascontiguousarray(data)), C.byref(L), eb, rel_eb, ptr(fl), ptr(qc), ptr(anchors),
        nthreads or nthreads_default())
    return qc, anchors


def spline_reconstruct(qc, anchors, shape, eb, rel_eb, flags, maxlevel=4, nthreads=None):
    nz, ny, nx = shape
    L = layout(nx, ny, nz)
    out = np.zeros(shape, anchors.dtype)
    fl = np.asarray(flags, np.uint8)
    getattr(lib(), f"po_spline_reconstruct_{sfx(anchors.dtype)}")(
        ptr(qc), ptr(anchors), C.byref(L), eb, rel_eb, ptr(fl), maxlevel, ptr(out),
        nthreads or nthreads_default())
    return out


def spline_delta_reconstruct(qc, old, eb, rel_eb, flags, maxlevel=4, nthreads=None):
    nz, ny, nx = old.shape
    L = layout(nx, ny, nz)
    out = np.zeros(old.shape, old.dtype)
    fl = np.asarray(flags, np.uint8)
    getattr(lib(), f"po_spline_delta_reconstruct_{sfx(old.dtype)}")(
        ptr(qc), C.byref(L), eb, rel_eb, ptr(fl), maxlevel, ptr(np.ascontiguousarray(old)), ptr(out),
        nthreads or nthreads_default())
    return out


def autotune(data: np.ndarray, rel_eb: float = 0.0):
    nz, ny, nx = data.shape
    L = layout(nx, ny, nz)
    scores = np.zeros(15, np.float64)
    flags = np.zeros(8, np.uint8)
    rc = getattr(lib(), f"po_autotune_{sfx(data.dtype)}")(
        ptr(np.ascontiguousarray(data)), C.byref(L), rel_eb, ptr(scores), ptr(flags))
    return scores, flags, rc


def pick_flags(scores) -> np.ndarray:
    s = np.asarray(scores, np.float64)
    flags = np.zeros(8, np.uint8)
    lib().po_pick_flags(ptr(s), ptr(flags))
    return flags


def bitshuffle(qc: np.ndarray, L: Layout, mode: int) -> np.ndarray:
    bp = np.zeros(L.host_size, np.uint8)
    lib().po_bitshuffle(ptr(qc), C.byref(L), mode, ptr(bp))
    return bp


def inverse_bitshuffle(bp: np.ndarray, L: Layout, mode: int, begin=None, end=None) -> np.ndarray:
    qc = np.zeros(L.code_len, np.int32)
    if begin is None:
        lib().po_inverse_bitshuffle(ptr(bp), C.byref(L), mode, ptr(qc))
    else:
        b = np.asarray(begin, np.int32)
        e = np.asarray(end, np.int32)
        lib().po_inverse_bitshuffle_window(ptr(bp), C.byref(L), mode, ptr(b), ptr(e), ptr(qc))
    return qc


def rze_encode_chunk(chunk: np.ndarray):
    chunk = np.ascontiguousarray(chunk, np.uint8)
    out = np.zeros(CS, np.uint8)
    n = lib().po_rze_encode_chunk(ptr(chunk), chunk.size, ptr(out))
    return None if n < 0 else out[:n].copy()


def rze_decode_chunk(enc: np.ndarray):
    enc = np.ascontiguousarray(enc, np.uint8)
    out = np.zeros(CS, np.uint8)
    n = lib().po_rze_decode_chunk(ptr(enc), enc.size, ptr(out))
    return None if n < 0 else out[:n].copy()


def lossless_encode(bp: np.ndarray, L: Layout):
    out = np.zeros(lib().po_stream_bound(C.byref(L)), np.uint8)
    cum = np.zeros(128, np.uint64)
    n = lib().po_lossless_encode(ptr(bp), C.byref(L), ptr(out), ptr(cum))
    return out[:n].copy(), cum


def lossless_decode(stream: np.ndarray, L: Layout):
    bp = np.zeros(L.host_size, np.uint8)
    rc = lib().po_lossless_decode(ptr(stream), stream.size, C.byref(L), ptr(bp))
    if rc != 0:
        raise ValueError(f"malformed stream ({rc})")
    return bp


def lossless_decode_window(stream: np.ndarray, L: Layout, begin, end, corrected=True, bp=None):
    if bp is None:
        bp = np.zeros(L.host_size, np.uint8)
    b = np.asarray(begin, np.int32)
    e = np.asarray(end, np.int32)
    n = lib().po_lossless_decode_window(ptr(stream), stream.size, C.byref(L), ptr(b), ptr(e),
                                        int(corrected), ptr(bp))
    if n < 0:
        raise ValueError(f"malformed stream ({n})")
    return bp, int(n)


def strategy_cost(lr: int, b: int) -> int:
    return int(lib().po_strategy_cost(lr, b))


def strategy_search(cum, target_cost: int):
    cum = np.ascontiguousarray(cum, np.uint64)
    ch = np.zeros(4, np.int32)
    bv = np.zeros(1, np.uint64)
    lib().po_strategy_search(ptr(cum), C.c_uint64(target_cost), ptr(ch), ptr(bv))
    return ch, int(bv[0])


def find_strategy(cum, begin, end, eb: float, target: float):
    cum = np.ascontiguousarray(cum, np.uint64)
    b = np.array(begin, np.int32)
    e = np.array(end, np.int32)
    used = np.zeros(1, np.int32)
    ml = lib().po_find_strategy(ptr(cum), ptr(b), ptr(e), eb, target, ptr(used))
    return b, e, ml, bool(used[0])


def compress(data: np.ndarray, eb: float, abs_mode=False, mode=0, nthreads=None):
    """Returns (stream bytes, flags). mode 0 = progressive (NB), 1 = SM."""
    nz, ny, nx = data.shape
    cap = lib().po_compress_bound(nx, ny, nz, data.dtype.itemsize)
    out = np.zeros(cap, np.uint8)
    flags = np.zeros(8, np.uint8)
    n = getattr(lib(), f"po_compress_{sfx(data.dtype)}")(
        ptr(np.ascontiguousarray(data)), nx, ny, nz, eb, int(abs_mode), mode, ptr(out), cap, ptr(flags),
        nthreads or nthreads_default())
    if n < 0:
        raise RuntimeError("oracle compress failed")
    return out[:n].copy(), flags


def decompress(stream: np.ndarray, shape, dtype, eb: float, abs_mode=False, nthreads=None):
    nz, ny, nx = shape
    out = np.zeros(shape, dtype)
    rc = getattr(lib(), f"po_decompress_{sfx(dtype)}")(
        ptr(stream), stream.size, nx, ny, nz, eb, int(abs_mode), ptr(out), nthreads or nthreads_default())
    if rc != 0:
        raise ValueError(f"oracle decompress failed ({rc})")
    return out


class Progressive:
    """Progressive retrieval state (begin/end per level, current output)."""

    def __init__(self, stream, shape, dtype, eb, abs_mode=False, corrected=True, nthreads=None):
        self.stream = np.ascontiguousarray(stream, np.uint8)
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.eb = eb
        self.abs_mode = abs_mode
        self.corrected = corrected
        self.begin = np.zeros(4, np.int32)
        self.end = np.zeros(4, np.int32)
        self.out = None
        self.rounds = 0
        self.nthreads = nthreads or nthreads_default()

    def retrieve(self, target: float):
        nz, ny, nx = self.shape
        new = np.zeros(self.shape, self.dtype)
        old = self.out if self.out is not None else np.zeros(self.shape, self.dtype)
        nb = np.zeros(1, np.int64)
        ml = getattr(lib(), f"po_progressive_round_{sfx(self.dtype)}")(
            ptr(self.stream), self.stream.size, nx, ny, nz, self.eb, int(self.abs_mode), target,
            int(self.rounds == 0), int(self.corrected), ptr(self.begin), ptr(self.end), ptr(old),
            ptr(new), ptr(nb), self.nthreads)
        if ml < 0:
            raise ValueError(f"progressive round failed ({ml})")
        self.out = new
        self.rounds += 1
        return new, int(ml), int(nb[0])
