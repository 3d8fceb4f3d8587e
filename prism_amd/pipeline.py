"""Host orchestration of the compress / decompress / progressive-retrieve path.

Follows the reference's Compressor flow (compressor.cu:120-190): extrema ->
auto-tune -> predict -> bit-plane shuffle -> lossless encode -> trailer, and
the inverse.  Every stage is a call into libprism_amd.so; torch is used only
to allocate device memory and to name the current HIP stream.  All stage
calls are stream-ordered; host synchronisation happens only where a host
value is needed (the value range of a REL-mode field, the stream length).
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import PA_F32, PA_F64, PA_MODE_NB, PA_MODE_SM, check, lib
from ._lib import layout as _new_layout

DTYPES = {torch.float32: PA_F32, torch.float64: PA_F64}


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


_LAYOUTS: dict = {}


def layout(nx: int, ny: int = 1, nz: int = 1):
    """The pa_layout of a shape, built once per shape and shared (every
    entry point takes it as const)."""
    key = (nx, ny, nz)
    L = _LAYOUTS.get(key)
    if L is None:
        L = _new_layout(nx, ny, nz)
        if len(_LAYOUTS) > 256:
            _LAYOUTS.clear()
        _LAYOUTS[key] = L
    return L


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


def _empty(nbytes: int, dtype=torch.uint8):
    item = dtype.itemsize
    return torch.empty(max(1, (int(nbytes) + item - 1) // item), dtype=dtype, device=_dev())


_TLS = threading.local()


def _host_read(t: torch.Tensor) -> np.ndarray:
    """Copies a small device tensor to the host through a pinned per-thread
    buffer and waits for the current stream; returns a fresh numpy copy."""
    nb = t.numel() * t.element_size()
    buf = getattr(_TLS, "pinned", None)
    if buf is None or buf.numel() < nb:
        buf = torch.empty(max(nb, 256), dtype=torch.uint8, pin_memory=True)
        _TLS.pinned = buf
    h = buf[:nb].view(t.dtype)
    h.copy_(t.reshape(-1), non_blocking=True)
    torch.cuda.current_stream().synchronize()
    return h.numpy().copy()


def shape_xyz(t: torch.Tensor):
    """(nx, ny, nz) of a [nz][ny][nx] (or lower-rank, x fastest) tensor."""
    s = list(t.shape)
    while len(s) < 3:
        s.insert(0, 1)
    if len(s) != 3:
        raise ValueError("fields are 1-D, 2-D or 3-D")
    return s[2], s[1], s[0]


@dataclass
class Compressed:
    stream: torch.Tensor          # uint8, exact length, on the device
    shape: tuple                  # (nx, ny, nz)
    dtype: torch.dtype
    eb_input: float               # rel_eb (REL) or eb (ABS)
    rel: bool
    progressive: bool
    range: float
    flags: list = field(default_factory=list)

    @property
    def nbytes(self) -> int:
        return int(self.stream.numel())

    def eb_abs(self) -> float:
        return self.eb_input * self.range if self.rel else self.eb_input


class StageTimer:
    """Per-stage device time from HIP events recorded on the current stream
    (the stream every stage launches on).  Durations are read after a
    synchronisation by `results()`."""

    def __init__(self):
        self.events = []  # (name, start, end)

    def __call__(self, name):
        timer = self

        class _Span:
            def __enter__(self_inner):
                self_inner.s = torch.cuda.Event(enable_timing=True)
                self_inner.e = torch.cuda.Event(enable_timing=True)
                self_inner.s.record()
                return self_inner

            def __exit__(self_inner, *exc):
                self_inner.e.record()
                timer.events.append((name, self_inner.s, self_inner.e))
                return False

        return _Span()

    def results(self):
        torch.cuda.synchronize()
        out = {}
        for name, s, e in self.events:
            out.setdefault(name, []).append(s.elapsed_time(e))
        return out


class _NoTimer:
    def __call__(self, name):
        return self

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


_NO_TIMER = _NoTimer()


def extrema(data: torch.Tensor):
    mm = torch.empty(4, dtype=torch.float64, device=data.device)
    check(lib().pa_extrema(_p(data), DTYPES[data.dtype], data.numel(), _p(mm), _stream()), "pa_extrema")
    return mm


QUALITY_FIELDS = ("linf", "linf_index", "sse", "min", "max", "mse", "psnr_db", "nrmse", "max_rel_err")


def quality(original: torch.Tensor, reconstruction: torch.Tensor) -> dict:
    """Error statistics of a reconstruction (statistic, analyze.hpp:5-70),
    computed on the device in one pass; one synchronisation to read them."""
    if original.shape != reconstruction.shape or original.dtype != reconstruction.dtype:
        raise _lib.PrismError("quality: original and reconstruction differ in shape or dtype")
    n = original.numel()
    out = torch.empty(9, dtype=torch.float64, device=original.device)
    ws = _empty(lib().pa_quality_workspace(n))
    check(lib().pa_quality(_p(original), _p(reconstruction), DTYPES[original.dtype], n, _p(out), _p(ws), ws.numel(),
                           _stream()), "pa_quality")
    vals = out.cpu().tolist()
    res = dict(zip(QUALITY_FIELDS, vals))
    res["linf_index"] = int(res["linf_index"])
    return res


def compress(data: torch.Tensor, eb: float, rel: bool = True, progressive: bool = True,
             return_parts: bool = False, timer=None):
    """Compresses a contiguous device field (T[nz][ny][nx]).

    The bound is resolved on the device (pa_level_bounds), so every stage is
    queued before the host waits once, at the end, for the stream length
    (the reference reads the range and the size table back in between)."""
    tm = timer or _NO_TIMER
    if not data.is_cuda or data.dtype not in DTYPES:
        raise ValueError("data must be a float32/float64 CUDA tensor")
    if not eb > 0:
        raise ValueError("error bound must be positive")
    data = data.contiguous()
    nx, ny, nz = shape_xyz(data)
    L = layout(nx, ny, nz)
    dt = DTYPES[data.dtype]
    st = _stream()
    dev = data.device
    V = C.c_void_p
    # the small per-call buffers in one allocation (byte offsets): bounds
    # double[16] at 0, the encoder's workspace at 128 (its error word first,
    # so that the one host read below takes it with the bounds), then min/max
    # double[4], scores double[15], flags uint8[8], cum int64[128], sizes
    # int64[2] and the tuner's workspace from e = 128 + encoder workspace
    tws_bytes = max(8, lib().pa_autotune_workspace(C.byref(L)))
    ews_bytes = lib().pa_encode_workspace(C.byref(L))
    e = 128 + (ews_bytes + 127) // 128 * 128
    small = torch.empty((e + 1216 + tws_bytes + 7) // 8, dtype=torch.float64, device=dev)
    sp = small.data_ptr()
    p_bounds, p_ews, p_mm, p_scores, p_flags, p_cum, p_sizes, p_tws = (
        V(sp), V(sp + 128), V(sp + e), V(sp + e + 32), V(sp + e + 152), V(sp + e + 160), V(sp + e + 1184),
        V(sp + e + 1216))
    if rel:  # extrema, auto-tuning and the level bounds share one launch sequence
        with tm("extrema"):
            check(lib().pa_extrema_autotune(_p(data), dt, C.byref(L), p_mm, float(eb), p_bounds, p_scores, p_flags,
                                            p_tws, tws_bytes, st), "pa_extrema_autotune")
    else:
        with tm("autotune"):
            check(lib().pa_autotune(_p(data), dt, C.byref(L), p_scores, p_flags, p_tws, tws_bytes, st), "pa_autotune")
        check(lib().pa_level_bounds(None, float(eb), dt, p_bounds, st), "pa_level_bounds")
    qc = torch.empty(L.code_len, dtype=torch.int32, device=dev)
    anchors = torch.empty(L.n4, dtype=data.dtype, device=dev)
    with tm("predict"):
        check(lib().pa_spline_construct_dev(_p(data), dt, C.byref(L), p_bounds, p_flags, _p(qc), _p(anchors), st),
              "pa_spline_construct_dev")
    bp = _empty(L.host_size)
    mode = PA_MODE_NB if progressive else PA_MODE_SM
    with tm("shuffle"):
        check(lib().pa_bitshuffle(_p(qc), C.byref(L), mode, _p(bp), st), "pa_bitshuffle")
    bound = lib().pa_compress_bound(C.byref(L), dt)
    out = _empty(bound)
    with tm("encode"):
        check(lib().pa_lossless_encode(_p(bp), C.byref(L), _p(out), p_cum, p_sizes, p_ews, ews_bytes, st),
              "pa_lossless_encode")
    with tm("merge"):  # the stream length goes to bounds[15] (int64 bits)
        check(lib().pa_compress_merge(_p(out), C.byref(L), dt, p_sizes, p_flags, V(sp + 8), p_cum, _p(anchors),
                                      V(sp + 120), st), "pa_compress_merge")
    # the one host read: stream length, status, range, the encoder's error word
    vals = _host_read(small[0:17])
    if vals[14] != 0:
        raise ValueError("error bound must be positive (constant field in REL mode?)")
    if vals[16:17].view(np.int64)[0] & 0xFFFFFFFF:
        raise _lib.PrismError("pa_lossless_encode: a chunk gave up waiting for the payload offsets before it")
    total = int(vals.view(np.int64)[15])
    rng = float(vals[1]) if rel else 1.0
    c = Compressed(stream=out[:total], shape=(nx, ny, nz), dtype=data.dtype, eb_input=eb, rel=rel,
                   progressive=progressive, range=rng, flags=[])
    if return_parts:
        e8 = e // 8
        return c, {"qc": qc, "anchors": anchors, "bp": bp, "cum": small[e8 + 20:e8 + 148].view(torch.int64),
                   "flags": small[e8 + 19:e8 + 20].view(torch.uint8), "scores": small[e8 + 4:e8 + 19], "layout": L}
    return c


class BatchCompressor:
    """Compresses a set of device fields of mixed shapes and dtypes with one
    fixed launch sequence (one launch per stage for all fields, captured in
    a HIP graph when `graph`); each stream equals compress() of that field.

    The plan is built for the shapes, dtypes, bound kinds and modes of
    `fields`; run() accepts new tensors and bounds of the same kinds.  The
    returned streams are views of a buffer that the next run() reuses."""

    def __init__(self, fields, ebs, rel: bool = True, progressive: bool = True, graph: bool = True,
                 autotune: bool = False):
        if not fields:
            raise ValueError("empty batch")
        self.rel, self.progressive = rel, progressive
        self.mode = PA_MODE_NB if progressive else PA_MODE_SM
        dev = fields[0].device
        self.nf = len(fields)
        self.shapes, self.dtypes = [], []
        offs, tot = [], 0
        for t in fields:
            if not t.is_cuda or t.dtype not in DTYPES:
                raise ValueError("fields must be float32/float64 CUDA tensors")
            nx, ny, nz = shape_xyz(t)
            L = layout(nx, ny, nz)
            self.shapes.append((nx, ny, nz))
            self.dtypes.append(t.dtype)
            offs.append(tot)
            tot += (lib().pa_compress_bound(C.byref(L), DTYPES[t.dtype]) + 255) // 256 * 256
        self.offs = offs
        self.outbuf = torch.empty(max(tot, 1), dtype=torch.uint8, device=dev)
        self.result = torch.zeros(4 * self.nf, dtype=torch.int64, device=dev)
        self._res_host = torch.empty(4 * self.nf, dtype=torch.int64, pin_memory=True)
        self._done = None
        self._keep = None
        desc = self._descs(fields, ebs)
        wsb = lib().pa_batch_workspace(desc, self.nf)
        if wsb == 0:
            raise _lib.PrismError("pa_batch_workspace: invalid field in the batch")
        self.ws = _empty(wsb)
        plan = C.c_void_p()
        check(lib().pa_batch_plan_create(desc, self.nf, _p(self.ws), self.ws.numel(), _p(self.result), int(graph),
                                         _stream(), C.byref(plan)), "pa_batch_plan_create")
        self.plan = plan
        self.graph = graph
        if autotune:
            check(lib().pa_batch_plan_autotune(plan, 3, _stream()), "pa_batch_plan_autotune")

    @property
    def params(self) -> dict:
        """Launch parameters in use and the autotuner's timings (ms)."""
        p = (C.c_int32 * 3)()
        t = (C.c_float * 9)()
        check(lib().pa_batch_plan_params(self.plan, p, t), "pa_batch_plan_params")
        return {"extrema_batches": p[0], "xcd_order": p[1], "encode_waves": p[2],
                "tuned_ms": {"extrema": list(t[0:4]), "spline_order": list(t[4:6]), "encode": list(t[6:9])}}

    def _descs(self, fields, ebs):
        if len(fields) != self.nf:
            raise ValueError("batch size differs from the plan")
        ebs = list(ebs) if hasattr(ebs, "__len__") else [ebs] * self.nf
        arr = (_lib.BatchFieldDesc * self.nf)()
        keep = []
        for i, t in enumerate(fields):
            t = t.contiguous()
            keep.append(t)
            nx, ny, nz = shape_xyz(t)
            d = arr[i]
            d.data, d.dtype, d.mode, d.rel = t.data_ptr(), DTYPES[t.dtype], self.mode, int(self.rel)
            d.eb, d.nx, d.ny, d.nz = float(ebs[i]), nx, ny, nz
            d.out = self.outbuf.data_ptr() + self.offs[i]
        self._keep = keep  # inputs stay alive while the plan points at them
        self._ebs = ebs
        return arr

    @property
    def launches(self) -> int:
        return int(lib().pa_batch_plan_launches(self.plan))

    def enqueue(self, fields=None, ebs=None):
        """Rebinds (when given new inputs) and launches; no synchronisation."""
        if fields is not None:
            desc = self._descs(fields, ebs if ebs is not None else self._ebs)
            check(lib().pa_batch_plan_rebind(self.plan, desc, self.nf, _stream()), "pa_batch_plan_rebind")
        check(lib().pa_batch_run(self.plan, _stream()), "pa_batch_run")
        self._res_host.copy_(self.result, non_blocking=True)
        self._done = torch.cuda.Event()
        self._done.record()

    @property
    def done(self) -> bool:
        """The last run (and its result copy) has completed; does not block."""
        return self._done is None or self._done.query()

    def results(self, strict: bool = True):
        """Per-field Compressed streams of the last run; waits for that run
        only.  A field whose bound is not positive (a constant field in REL
        mode) raises, or gives None when not `strict`."""
        if self._done is not None:
            self._done.synchronize()
        r = self._res_host.view(-1, 4)
        out = []
        for i in range(self.nf):
            total, status = int(r[i, 0]), int(r[i, 1])
            if status != 0:
                if strict:
                    raise _lib.PrismError(f"batch field {i}: non-positive error bound (constant field in REL mode?)")
                out.append(None)
                continue
            rng = float(r[i, 2:3].view(torch.float64).item())
            s = self.outbuf[self.offs[i]:self.offs[i] + total]
            out.append(Compressed(stream=s, shape=self.shapes[i], dtype=self.dtypes[i], eb_input=float(self._ebs[i]),
                                  rel=self.rel, progressive=self.progressive, range=rng if self.rel else 1.0))
        return out

    def run(self, fields=None, ebs=None):
        self.enqueue(fields, ebs)
        return self.results()

    def __del__(self):
        plan = getattr(self, "plan", None)
        if plan:
            try:
                if self._done is not None:
                    self._done.synchronize()  # no run of the plan may be in flight
                lib().pa_batch_plan_destroy(plan)
            except Exception:
                pass
            self.plan = None


def _trailer(c: Compressed, L):
    item = 4 if c.dtype == torch.float32 else 8
    tsize = L.trailer_fixed + L.n4 * item
    t0 = c.nbytes - tsize
    if t0 < 8 + 2 * L.nchunks or t0 % 8:
        raise _lib.PrismError("stream too short or misaligned for this shape")
    s = c.stream
    flags = s[t0:t0 + 8]
    cum = s[t0 + 16:t0 + 1040].view(torch.int64)
    anchors = s[t0 + 1040:t0 + 1040 + L.n4 * item].view(c.dtype)
    return t0, flags, cum, anchors


def stream_range(c: Compressed) -> float:
    """The value range stored in the trailer (compressor.cu:146, 204)."""
    nx, ny, nz = c.shape
    L = layout(nx, ny, nz)
    t0, _, _, _ = _trailer(c, L)
    return float(c.stream[t0 + 8:t0 + 16].view(torch.float64).item())


def _decoder_error(dws: torch.Tensor) -> int:
    """The decoder's error word (first int32 of its workspace); waits for the
    current stream."""
    return int(_host_read(dws[0:4].view(torch.int32))[0])


def decompress(c: Compressed, check_stream: bool = True) -> torch.Tensor:
    """Full decompression of a plain (SM) stream, decompress_pipeline
    (compressor.cu:153-164).  With `check_stream` (default) the call waits
    for the device and raises PrismError when the decoder found the stream
    malformed (a size-table entry pointing outside the stream: the chunks it
    names decode to zeros); without it nothing is synchronised."""
    nx, ny, nz = c.shape
    L = layout(nx, ny, nz)
    dt = DTYPES[c.dtype]
    st = _stream()
    t0, flags, cum, anchors = _trailer(c, L)
    rng = stream_range(c)
    eb_abs = c.eb_input * rng if c.rel else c.eb_input
    rel_eb = c.eb_input if c.rel else 0.0
    bp = _empty(L.host_size)
    ws = _empty(lib().pa_decode_workspace(C.byref(L)))
    check(lib().pa_lossless_decode(_p(c.stream), c.nbytes, C.byref(L), _p(bp), _p(ws), ws.numel(), st), "pa_lossless_decode")
    qc = torch.empty(L.code_len, dtype=torch.int32, device=c.stream.device)
    mode = PA_MODE_NB if c.progressive else PA_MODE_SM
    check(lib().pa_inverse_bitshuffle(_p(bp), C.byref(L), mode, _p(qc), st), "pa_inverse_bitshuffle")
    out = torch.empty((nz, ny, nx), dtype=c.dtype, device=c.stream.device)
    check(lib().pa_spline_reconstruct(_p(qc), _p(anchors), dt, C.byref(L), eb_abs, rel_eb, _p(flags), None,
                                      _p(out), st), "pa_spline_reconstruct")
    if check_stream:
        err = _decoder_error(ws)
        if err:
            raise _lib.PrismError(f"malformed compressed stream (decoder status {err})")
    return out


class Retriever:
    """Progressive retrieval of one compressed field, keeping its state
    (begin/end per level, decoded bit-planes, current reconstruction)
    resident on the device; each retrieve() refines from the current level
    (decompress_progressive_pipeline, compressor.cu:166-190)."""

    def __init__(self, c: Compressed):
        if not c.progressive:
            raise ValueError("progressive retrieval needs a stream compressed with progressive=True")
        self.c = c
        nx, ny, nz = c.shape
        self.L = layout(nx, ny, nz)
        self.dt = DTYPES[c.dtype]
        dev = c.stream.device
        # the range compress() returns is the one the trailer records
        self.range = c.range if c.range is not None else stream_range(c)
        self.eb_abs = c.eb_input * self.range if c.rel else c.eb_input
        self.rel_eb = c.eb_input if c.rel else 0.0
        # state (begin[4] end[4] maxlevel narrow, int32[16]) and the byte
        # counter (int64) in one zero-filled block; views on demand
        self._small = torch.zeros(18, dtype=torch.int32, device=dev)
        # the strategy search's best key and arrival count (zero before first use)
        self.sws = torch.zeros(lib().pa_strategy_workspace(), dtype=torch.uint8, device=dev)
        self._bufs = None  # bit-planes, codes and decoder workspace, allocated on first use
        self.out = None
        self.rounds = 0
        # argument values of the per-round calls, from base addresses (the
        # trailer layout of _trailer: flags at t0, cum at t0 + 16, anchors at
        # t0 + 1040)
        item = 4 if c.dtype == torch.float32 else 8
        self.t0 = c.nbytes - (self.L.trailer_fixed + self.L.n4 * item)
        if self.t0 < 8 + 2 * self.L.nchunks or self.t0 % 8:
            raise _lib.PrismError("stream too short or misaligned for this shape")
        sb, sp = c.stream.data_ptr(), self._small.data_ptr()
        V = C.c_void_p
        self._a = dict(cum=V(sb + self.t0 + 16), begin=V(sp), end=V(sp + 16), maxlevel=V(sp + 32),
                       sws=_p(self.sws), stream=V(sb), bytes_read=V(sp + 64), narrow=V(sp + 36),
                       anchors=V(sb + self.t0 + 1040), flags=V(sb + self.t0), L=C.byref(self.L))

    def _buffers(self):
        """(bp, qc, dws), allocated at first use: the first round's strategy
        search is already queued by then."""
        if self._bufs is None:
            dev = self._small.device
            bp = _empty(self.L.host_size)
            qc = torch.empty(self.L.code_len, dtype=torch.int32, device=dev)
            dws = _empty(lib().pa_decode_workspace(C.byref(self.L)))
            self._bufs = (bp, qc, dws)
            self._a.update(bp=_p(bp), qc=_p(qc), dws=_p(dws))
        return self._bufs

    @property
    def bp(self):
        return self._buffers()[0]

    @property
    def qc(self):
        return self._buffers()[1]

    @property
    def dws(self):
        return self._buffers()[2]

    # views of the device state and of the trailer
    @property
    def state(self):
        return self._small[0:16]

    @property
    def begin(self):
        return self._small[0:4]

    @property
    def end(self):
        return self._small[4:8]

    @property
    def maxlevel(self):
        return self._small[8:9]

    @property
    def narrow(self):  # this round's codes: 0 int32 hand-off, 1 int16 hand-off, 2 read from the bit-planes
        return self._small[9:10]

    @property
    def bytes_read(self):
        return self._small[16:18].view(torch.int64)

    @property
    def flags(self):
        return _trailer(self.c, self.L)[1]

    @property
    def cum(self):
        return _trailer(self.c, self.L)[2]

    @property
    def anchors(self):
        return _trailer(self.c, self.L)[3]

    def error(self) -> int:
        """The decoder's error bits for this field so far (0: every chunk
        read was well formed); waits for the current stream.  The bits are
        sticky: once a round has met a malformed chunk, every later round of
        the field reports it."""
        if self._bufs is None:
            return 0
        return _decoder_error(self._bufs[2])

    def check(self) -> None:
        """Raises PrismError if a round of this field met a malformed chunk."""
        err = self.error()
        if err:
            raise _lib.PrismError(f"malformed compressed stream (decoder status {err})")

    def retrieve(self, target: float, timer=None, check_stream: bool = True) -> torch.Tensor:
        """Refines to the target error bound (relative in REL mode).  With
        `check_stream` (default) the call waits for the round and raises
        PrismError if the decoder met a malformed chunk; without it the round
        is only queued, and check() reports the error later."""
        tm = timer or _NO_TIMER
        st = _stream()
        a = self._a
        lb = lib()
        target_abs = target * self.range if self.c.rel else target
        with tm("strategy"):
            check(lb.pa_find_strategy(a["cum"], a["begin"], a["end"], a["maxlevel"], self.eb_abs, target_abs,
                                      a["sws"], self.sws.numel(), st), "pa_find_strategy")
        dws = self._buffers()[2]
        with tm("decode"):
            check(lb.pa_lossless_decode_window(a["stream"], self.c.nbytes, a["L"], a["begin"], a["end"], a["bp"],
                                               a["bytes_read"], int(self.rounds > 0), a["dws"], dws.numel(), st),
                  "pa_lossless_decode_window")
        with tm("unshuffle"):
            # no codes written when the reconstruction reads the planes itself
            check(lb.pa_inverse_bitshuffle_window_deferred(a["bp"], a["L"], a["begin"], a["end"], a["qc"],
                                                           a["narrow"], st), "pa_inverse_bitshuffle_window_deferred")
        nx, ny, nz = self.c.shape
        new = torch.empty((nz, ny, nx), dtype=self.c.dtype, device=self._small.device)
        with tm("reconstruct"):
            self._reconstruct(new, st)
        self.out = new
        self.rounds += 1
        if check_stream:
            self.check()
        return new

    def job(self, target: float, out: torch.Tensor, desc) -> None:
        """Fills a pa_retrieve_job for one round of this field (batched
        retrieval); the round is committed by `commit(out)` once enqueued."""
        c = self.c
        nx, ny, nz = c.shape
        desc.stream, desc.stream_len, desc.dtype = c.stream.data_ptr(), c.nbytes, self.dt
        desc.nx, desc.ny, desc.nz = nx, ny, nz
        desc.eb, desc.rel_eb = self.eb_abs, self.rel_eb
        desc.target = target * self.range if c.rel else target
        sp = self._small.data_ptr()
        desc.state, desc.bp, desc.qc = sp, self.bp.data_ptr(), self.qc.data_ptr()
        desc.dws, desc.bytes_read = self.dws.data_ptr(), sp + 64
        desc.round = self.rounds
        desc.old = self.out.data_ptr() if self.out is not None else None
        desc.out = out.data_ptr()

    def commit(self, out: torch.Tensor) -> None:
        self.out = out
        self.rounds += 1

    def _reconstruct(self, new, st):
        a = self._a
        first = self.out is None
        check(lib().pa_spline_reconstruct_planes(a["qc"], a["narrow"], a["bp"], a["begin"], a["end"],
                                                 a["anchors"] if first else None, None if first else _p(self.out),
                                                 self.dt, a["L"], self.eb_abs, self.rel_eb, a["flags"], a["maxlevel"],
                                                 _p(new), st), "pa_spline_reconstruct_planes")


class RetrievalBatch:
    """One enqueued batch of retrieval rounds: the new reconstructions, the
    jobs' decoder error words (copied to pinned host memory behind the
    batch) and the buffers that must outlive the device work."""

    def __init__(self, outs, keep, errors=None):
        self.outs = outs
        self._keep = keep
        self._errors = errors

    def errors(self):
        """Per job, the field's decoder error bits after this round (0: well
        formed).  Valid once the batch's work has completed (an event
        recorded after it has fired, or after a synchronisation)."""
        if self._errors is None:
            return [0] * len(self.outs)
        return [int(v) for v in self._errors.tolist()]


def retrieve_batch(items, pinned: bool = False) -> RetrievalBatch:
    """One retrieval round for each (Retriever, target) of `items` (distinct
    fields) with one fixed launch sequence; each result equals
    Retriever.retrieve(target) on that field.  With `pinned`, the job table
    goes through pinned host memory asynchronously (the returned object then
    holds it until the caller has synchronised); otherwise the call
    synchronises the current stream before its upload."""
    nj = len(items)
    if nj == 0:
        return RetrievalBatch([], None)
    if len({id(r) for r, _ in items}) != nj:
        raise ValueError("a batch takes one round per field")
    arr = (_lib.RetrieveJobDesc * nj)()
    outs = []
    for k, (r, target) in enumerate(items):
        nx, ny, nz = r.c.shape
        new = torch.empty((nz, ny, nx), dtype=r.c.dtype, device=r.qc.device)
        r.job(float(target), new, arr[k])
        outs.append(new)
    ws = _empty(lib().pa_retrieve_workspace(nj))
    h = None
    tb = 0
    if pinned:
        tb = lib().pa_retrieve_table_bytes(nj)
        h = torch.empty(tb, dtype=torch.uint8, pin_memory=True)
    check(lib().pa_retrieve_batch(arr, nj, _p(h), tb, _p(ws), ws.numel(), _stream()), "pa_retrieve_batch")
    # each job's error word, gathered by one kernel and copied to the host
    # behind the batch (read by errors() once the batch has completed)
    errs = torch.cat([r.dws[0:4].view(torch.int32) for r, _ in items])
    errs_h = torch.empty(nj, dtype=torch.int32, pin_memory=True)
    errs_h.copy_(errs, non_blocking=True)
    # the previous reconstructions the batch reads (job.old) stay alive until
    # the caller has synchronised, even if commit() drops the retriever's
    # own reference to them
    olds = [r.out for r, _ in items]
    for (r, _), new in zip(items, outs):
        r.commit(new)
    return RetrievalBatch(outs, (ws, h, olds, errs), errs_h)
