"""The allocation-free decode (pa_lossless_decode_noalloc) against the
workspace decode (pa_lossless_decode), through the C ABI: bit-planes and
error word equal, bit for bit, over consecutive calls that share the one
scratch array -- a large field, a malformed stream, then a smaller field
(whose error word must read 0 again) -- and fields above the scratch limit
rejected before any device work."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


def P(t):
    return C.c_void_p(t.data_ptr())


def S():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.fixture(scope="module")
def lib():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from prism_amd import _lib
    _lib.lib()
    return _lib


def _encode(_lib, L, bp):
    out = torch.zeros(_lib.lib().pa_compress_bound(C.byref(L), 8), dtype=torch.uint8, device="cuda")
    cum = torch.zeros(128, dtype=torch.int64, device="cuda")
    sz = torch.zeros(2, dtype=torch.int64, device="cuda")
    ws = torch.empty(_lib.lib().pa_encode_workspace(C.byref(L)), dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().pa_lossless_encode(P(bp), C.byref(L), P(out), P(cum), P(sz), P(ws), ws.numel(), S()))
    return out


def _both(_lib, L, stream):
    """(bit-planes, error word) of the workspace decode and of the
    allocation-free decode of the same stream"""
    lib = _lib.lib()
    ref = torch.full((L.host_size,), 0x5A, dtype=torch.uint8, device="cuda")
    ws = torch.empty(lib.pa_decode_workspace(C.byref(L)), dtype=torch.uint8, device="cuda")
    _lib.check(lib.pa_lossless_decode(P(stream), stream.numel(), C.byref(L), P(ref), P(ws), ws.numel(), S()),
               "pa_lossless_decode")
    got = torch.full((L.host_size,), 0xA5, dtype=torch.uint8, device="cuda")
    err = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    _lib.check(lib.pa_lossless_decode_noalloc(P(stream), stream.numel(), C.byref(L), P(got), P(err), S()),
               "pa_lossless_decode_noalloc")
    torch.cuda.synchronize()
    return ref.cpu(), int(ws[:4].view(torch.int32).cpu()[0]), got.cpu(), int(err.cpu()[0])


def test_noalloc_decode_equals_workspace_decode(lib):
    rng = np.random.default_rng(11)
    # (nx, ny, nz): a field of many chunks first, then smaller ones
    cases = [((384, 256, 64), "valid"), ((130, 70, 33), "corrupt"), ((64, 40, 20), "valid"), ((300, 1, 1), "valid")]
    for (nx, ny, nz), kind in cases:
        announce("noalloc decode", (nx, ny, nz), kind)
        L = lib.layout(nx, ny, nz)
        assert lib.lib().pa_decode_workspace(C.byref(L)) <= lib.lib().pa_noalloc_scratch_bytes()
        p = np.repeat(rng.choice([0.0, 0.02, 0.5, 1.0], (L.host_size + 1023) // 1024), 1024)[:L.host_size]
        bp_np = np.where(rng.random(L.host_size) < p, rng.integers(1, 256, L.host_size), 0).astype(np.uint8)
        bp = torch.from_numpy(bp_np).cuda()
        stream = _encode(lib, L, bp)
        if kind == "corrupt":
            # the first chunk's size entry points past the stream's payload
            stream[8:10] = torch.tensor([0xFF, 0x7F], dtype=torch.uint8)
        ref, ref_err, got, got_err = _both(lib, L, stream)
        assert torch.equal(got, ref)
        assert got_err == ref_err
        if kind == "valid":
            assert got_err == 0
            assert np.array_equal(got.numpy(), bp_np)
        else:
            assert got_err != 0


def test_noalloc_decode_rejects_fields_above_the_scratch(lib):
    announce("noalloc decode size limit")
    L = lib.layout(2048, 2048, 128)  # 2^29 elements: more chunks than the scratch holds
    need = lib.lib().pa_decode_workspace(C.byref(L))
    assert need > lib.lib().pa_noalloc_scratch_bytes()
    buf = torch.zeros(64, dtype=torch.uint8, device="cuda")
    err = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = lib.lib().pa_lossless_decode_noalloc(P(buf), buf.numel(), C.byref(L), P(buf), P(err), S())
    assert rc == lib.PA_ERR_WORKSPACE
