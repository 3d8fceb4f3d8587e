"""Parity of the HIP path with the CPU oracle, stage by stage and end to end.

Integer and byte outputs (codes, bit-planes, streams, strategy) must be
bit-exact; floating-point outputs (reconstructions, tuner scores) must be
bit-exact too, because both sides evaluate the same IEEE operations in the
same order with no FMA contraction.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pa():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import prism_amd
    from prism_amd import _lib, pipeline
    prism_amd.lib()
    return _lib, pipeline


def smooth(shape, dtype, seed=0, noise=1e-3):
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = (np.sin(x / 5.3 + rng.random()) * np.cos(y / 7.1) + 0.5 * np.sin(z / 3.7 + 1.0) +
         0.2 * np.cos((x + y) / 11.0) + rng.normal(0, noise, shape))
    return np.ascontiguousarray(f.astype(dtype))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def P(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def S():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


# (nz, ny, nx); 1-D shapes span one and several 256-tile workgroups of the
# line-field kernel, 2-D shapes one and several 4-tile workgroups of the
# plane kernel, with partial last tiles; (6, 1, 70) is a 3-D field one row
# high (the tile kernel with a single y)
SHAPES = [(33, 47, 29), (16, 16, 16), (1, 1, 1000), (1, 40, 70), (35, 18, 20), (64, 64, 64), (17, 17, 17),
          (1, 1, 5), (1, 1, 4097), (1, 1, 12345), (1, 2, 300), (1, 17, 33), (1, 130, 75), (6, 1, 70)]
FLAGSETS = [[0] * 8, [1] * 4 + [0] * 4, [0] * 4 + [1] * 4, [1, 0, 1, 0, 0, 1, 0, 1]]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_extrema(pa, dtype):
    _lib, pipe = pa
    rng = np.random.default_rng(1)
    for n in (1, 3, 1000, (1 << 20) + 5):
        announce("extrema", dtype.__name__, n)
        a = rng.normal(0, 10, n).astype(dtype)
        mm = host(pipe.extrema(dev(a)))
        assert mm[0] == float(a.min()) and mm[1] == float(a.max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(33, 47, 29), (65, 101, 129)])
def test_extrema_planted(pa, dtype, shape):
    """Extremes at the first element, in the scalar tail (n is not a whole
    number of 16-byte vectors) and in the middle, over all-negative and
    mixed-sign fields: the order-preserving key image flips every bit of a
    negative value, so a sign slip there shows up as a wrong min or max.
    Both the standalone entry and the one beside the tuner."""
    _lib, pipe = pa
    nz, ny, nx = shape
    n = nx * ny * nz
    assert n % (16 // np.dtype(dtype).itemsize) != 0
    L = _lib.layout(nx, ny, nz)
    dt = _lib.PA_F32 if dtype == np.float32 else _lib.PA_F64
    ws = torch.empty(max(8, _lib.lib().pa_autotune_workspace(C.byref(L))), dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(5)
    big = 1e30 if dtype == np.float32 else 1e300
    for case, (imin, imax) in enumerate([(0, n - 1), (n - 1, 0), (n // 2, n // 3)]):
        for offset in (-50.0, 0.0):
            announce("extrema planted", shape, dtype.__name__, imin, imax, offset)
            a = (rng.uniform(-1, 1, n) + offset).astype(dtype)
            a[imin] = -big if offset == 0.0 else offset - 10
            a[imax] = big if offset == 0.0 else offset + 10
            mm = host(pipe.extrema(dev(a)))
            assert mm[0] == float(a.min()) and mm[1] == float(a.max())
            mm2 = torch.empty(4, dtype=torch.float64, device="cuda")
            sc = torch.empty(15, dtype=torch.float64, device="cuda")
            fl = torch.empty(8, dtype=torch.uint8, device="cuda")
            bd = torch.zeros(16, dtype=torch.float64, device="cuda")
            _lib.check(_lib.lib().pa_extrema_autotune(P(dev(a)), dt, C.byref(L), P(mm2), 1e-4, P(bd), P(sc), P(fl),
                                                      P(ws), ws.numel(), S()))
            mmh = host(mm2)
            assert mmh[0] == float(a.min()) and mmh[1] == float(a.max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", SHAPES)
def test_predictor_parity(pa, oracle, dtype, shape):
    _lib, pipe = pa
    nz, ny, nx = shape
    data = smooth(shape, dtype, seed=nx)
    L = _lib.layout(nx, ny, nz)
    dt = _lib.PA_F32 if dtype == np.float32 else _lib.PA_F64
    for flags in FLAGSETS:
        for rel in (1e-2, 1e-5):
            announce("predictor", shape, dtype.__name__, flags, rel)
            rg = float(data.max()) - float(data.min())
            eb = rel * rg
            qc_o, an_o = oracle.spline_construct(data, eb, rel, flags)
            d = dev(data)
            fl = dev(np.array(flags, np.uint8))
            qc = torch.empty(L.code_len, dtype=torch.int32, device="cuda")
            an = torch.empty(L.n4, dtype=d.dtype, device="cuda")
            _lib.check(_lib.lib().pa_spline_construct(P(d), dt, C.byref(L), eb, rel, P(fl), P(qc), P(an), S()))
            assert np.array_equal(host(qc), qc_o)
            assert np.array_equal(host(an), an_o)
            rec_o = oracle.spline_reconstruct(qc_o, an_o, shape, eb, rel, flags)
            out = torch.empty(shape, dtype=d.dtype, device="cuda")
            _lib.check(_lib.lib().pa_spline_reconstruct(P(qc), P(an), dt, C.byref(L), eb, rel, P(fl), None,
                                                         P(out), S()))
            assert np.array_equal(host(out).view(np.uint8), rec_o.view(np.uint8))
            # partial level ranges: anchored form and delta form
            half = qc_o // 3
            old = rec_o * dtype(0.5)
            half_d, old_d = dev(half), dev(old)  # keep alive until the kernels have run
            for maxlevel in (0, 1, 2, 3):
                ml = torch.tensor([maxlevel], dtype=torch.int32, device="cuda")
                part_o = oracle.spline_reconstruct(qc_o, an_o, shape, eb, rel, flags, maxlevel=maxlevel)
                out1 = torch.empty(shape, dtype=d.dtype, device="cuda")
                _lib.check(_lib.lib().pa_spline_reconstruct(P(qc), P(an), dt, C.byref(L), eb, rel, P(fl), P(ml),
                                                             P(out1), S()))
                assert np.array_equal(host(out1).view(np.uint8), part_o.view(np.uint8))
                dl_o = oracle.spline_delta_reconstruct(half, old, eb, rel, flags, maxlevel=maxlevel)
                out2 = torch.empty(shape, dtype=d.dtype, device="cuda")
                _lib.check(_lib.lib().pa_spline_delta_reconstruct(P(half_d), dt, C.byref(L), eb, rel, P(fl),
                                                                   P(ml), P(old_d), P(out2), S()))
                assert np.array_equal(host(out2).view(np.uint8), dl_o.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(40, 40, 40), (20, 70, 33), (35, 300, 17), (17, 17, 17), (1, 64, 64)])
def test_tuner_parity(pa, oracle, dtype, shape):
    _lib, pipe = pa
    nz, ny, nx = shape
    announce("tuner", shape, dtype.__name__)
    data = smooth(shape, dtype, seed=3, noise=0.05)
    sc_o, fl_o, rc = oracle.autotune(data)
    L = _lib.layout(nx, ny, nz)
    dt = _lib.PA_F32 if dtype == np.float32 else _lib.PA_F64
    d = dev(data)
    fl = torch.empty(8, dtype=torch.uint8, device="cuda")
    sc = torch.empty(15, dtype=torch.float64, device="cuda")
    ws = torch.empty(max(8, _lib.lib().pa_autotune_workspace(C.byref(L))), dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().pa_autotune(P(d), dt, C.byref(L), P(sc), P(fl), P(ws), ws.numel(), S()))
    assert np.array_equal(host(sc).view(np.uint64), sc_o.view(np.uint64))
    assert list(host(fl)) == list(fl_o)
    # the extrema and the tuner in one launch: the same scores and flags, and
    # the exact min / max (a field too small for tuner samples takes the two
    # separate calls inside)
    announce("extrema + tuner", shape, dtype.__name__)
    mm = torch.empty(4, dtype=torch.float64, device="cuda")
    fl2 = torch.full((8,), 7, dtype=torch.uint8, device="cuda")
    sc2 = torch.empty(15, dtype=torch.float64, device="cuda")
    bd = torch.zeros(16, dtype=torch.float64, device="cuda")  # slots of the other dtype stay 0
    _lib.check(_lib.lib().pa_extrema_autotune(P(d), dt, C.byref(L), P(mm), 1e-4, P(bd), P(sc2), P(fl2), P(ws),
                                              ws.numel(), S()))
    assert np.array_equal(host(sc2).view(np.uint64), sc_o.view(np.uint64))
    assert list(host(fl2)) == list(fl_o)
    mmh = host(mm)
    assert mmh[0] == float(data.min()) and mmh[1] == float(data.max())
    # the level bounds it appends equal pa_level_bounds on the same min / max
    bd2 = torch.zeros(16, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().pa_level_bounds(P(mm), 1e-4, dt, P(bd2), S()))
    assert np.array_equal(host(bd)[:16].view(np.uint64), host(bd2)[:16].view(np.uint64))


def _codes_for(oracle, shape, dtype, rel=1e-4, seed=0):
    data = smooth(shape, dtype, seed=seed)
    rg = float(data.max()) - float(data.min())
    qc, _ = oracle.spline_construct(data, rel * rg, rel, [1, 0, 1, 0, 0, 1, 0, 1])
    return qc


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("shape", SHAPES + [(100, 30, 1), (300, 1, 1)])
def test_shuffle_parity(pa, oracle, shape, mode):
    _lib, pipe = pa
    nz, ny, nx = shape
    announce("shuffle", shape, mode)
    L = _lib.layout(nx, ny, nz)
    Lo = oracle.layout(nx, ny, nz)
    qc_o = _codes_for(oracle, shape, np.float64)
    bp_o = oracle.bitshuffle(qc_o, Lo, mode)
    qc = dev(qc_o)
    bp = torch.full((L.host_size,), 0xCD, dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().pa_bitshuffle(P(qc), C.byref(L), mode, P(bp), S()))
    assert np.array_equal(host(bp), bp_o)
    back = torch.zeros(L.code_len, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().pa_inverse_bitshuffle(P(bp), C.byref(L), mode, P(back), S()))
    assert np.array_equal(host(back)[L.n4:L.real_end], qc_o[L.n4:L.real_end])
    for b, e in (([0] * 4, [5, 9, 32, 1]), ([3, 0, 7, 31], [20, 32, 7, 32])):
        w_o = oracle.inverse_bitshuffle(bp_o, Lo, mode, b, e)
        be = dev(np.array(b + e, np.int32))
        w = torch.zeros(L.code_len, dtype=torch.int32, device="cuda")
        _lib.check(_lib.lib().pa_inverse_bitshuffle_window(P(bp), C.byref(L), mode, P(be[0:4]), P(be[4:8]), P(w),
                                                            S()))
        assert np.array_equal(host(w)[L.n4:L.real_end], w_o[L.n4:L.real_end])


def _gpu_encode(_lib, L, bp_np):
    bp = dev(bp_np)
    bound = _lib.lib().pa_compress_bound(C.byref(L), 8)
    out = torch.zeros(bound, dtype=torch.uint8, device="cuda")
    cum = torch.zeros(128, dtype=torch.int64, device="cuda")
    sz = torch.zeros(2, dtype=torch.int64, device="cuda")
    ws = torch.empty(_lib.lib().pa_encode_workspace(C.byref(L)), dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().pa_lossless_encode(P(bp), C.byref(L), P(out), P(cum), P(sz), P(ws), ws.numel(), S()))
    n = int(host(sz)[0])
    return out, host(out)[:n], host(cum).view(np.uint64)


@pytest.mark.parametrize("shape", [(33, 47, 29), (64, 64, 64), (300, 1, 1), (1, 100, 70), (128, 128, 128)])
@pytest.mark.parametrize("kind", ["codes", "sparse", "dense", "zeros", "full", "blocky"])
def test_codec_parity(pa, oracle, shape, kind):
    """Chunk codec against the oracle.  `full` (no zero byte) stores every
    chunk raw; `blocky` gives each 1 KiB step of a chunk its own density --
    empty, sparse, dense or full -- so the zero-eliminated bytes skip whole
    steps and carry partial words across them."""
    _lib, pipe = pa
    nz, ny, nx = shape
    announce("codec", shape, kind)
    L = _lib.layout(nx, ny, nz)
    Lo = oracle.layout(nx, ny, nz)
    rng = np.random.default_rng(2)
    if kind == "codes":
        bp_np = oracle.bitshuffle(_codes_for(oracle, shape, np.float32), Lo, 0)
    elif kind == "sparse":
        bp_np = np.where(rng.random(L.host_size) < 0.03, rng.integers(1, 256, L.host_size), 0).astype(np.uint8)
    elif kind == "dense":
        bp_np = np.where(rng.random(L.host_size) < 0.6, rng.integers(1, 256, L.host_size), 0).astype(np.uint8)
    elif kind == "full":
        bp_np = rng.integers(1, 256, L.host_size).astype(np.uint8)
    elif kind == "blocky":
        nblk = (L.host_size + 1023) // 1024
        p = np.repeat(rng.choice([0.0, 0.01, 0.5, 1.0], nblk), 1024)[:L.host_size]
        bp_np = np.where(rng.random(L.host_size) < p, rng.integers(1, 256, L.host_size), 0).astype(np.uint8)
    else:
        bp_np = np.zeros(L.host_size, np.uint8)
    s_o, cum_o = oracle.lossless_encode(bp_np, Lo)
    out, s_g, cum_g = _gpu_encode(_lib, L, bp_np)
    pad = (-s_g.size) % 8
    assert s_g.size + pad == s_o.size
    assert np.array_equal(s_g, s_o[:s_g.size])
    assert np.array_equal(cum_g, cum_o)
    bp = torch.full((L.host_size,), 0x5A, dtype=torch.uint8, device="cuda")
    ws = torch.empty(_lib.lib().pa_decode_workspace(C.byref(L)), dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().pa_lossless_decode(P(out), out.numel(), C.byref(L), P(bp), P(ws), ws.numel(), S()))
    assert np.array_equal(host(bp), bp_np)
    assert int(host(ws[:4].view(torch.int32))[0]) == 0
    # windowed decode: every device byte of the requested planes is restored
    b, e = [1, 0, 6, 30], [9, 32, 6, 32]
    be = dev(np.array(b + e, np.int32))
    bp2 = torch.zeros(L.host_size, dtype=torch.uint8, device="cuda")
    nread = torch.zeros(1, dtype=torch.int64, device="cuda")
    _lib.check(_lib.lib().pa_lossless_decode_window(P(out), out.numel(), C.byref(L), P(be[0:4]), P(be[4:8]), P(bp2), P(nread),
                                                     0, P(ws), ws.numel(), S()))
    got = host(bp2)
    ref, nb = oracle.lossless_decode_window(s_o, Lo, b, e, corrected=True)
    assert int(host(nread)[0]) == nb
    for lr in range(4):
        lv = 3 - lr
        if b[lr] < e[lr]:
            lo = L.dev_base[lv] + b[lr] * L.dev_stride[lv]
            hi = L.dev_base[lv] + e[lr] * L.dev_stride[lv]
            assert np.array_equal(got[lo:hi], bp_np[lo:hi])


def test_codec_parity_many_chunks(pa, oracle):
    """A stream of more than 16,384 chunks (512 x 512 x 300 codes: 19,200
    chunks, 10,000+ encoder workgroups, so the look-back runs over several
    rounds of workgroups and the decoder's size scan over several windows):
    stream, size table and cumulative table byte-equal to the oracle's,
    full and windowed decode equal to the planes."""
    _lib, pipe = pa
    nx, ny, nz = 512, 512, 300
    announce("codec many chunks", (nz, ny, nx))
    L = _lib.layout(nx, ny, nz)
    Lo = oracle.layout(nx, ny, nz)
    assert L.nchunks > 16384
    rng = np.random.default_rng(7)
    # each 1 KiB block empty, sparse, half, dense or full (uint8 draws: fast
    # at 314 MB); a byte survives where its draw is below the block's level
    nblk = (L.host_size + 1023) // 1024
    level = np.repeat(rng.choice(np.array([0, 1, 128, 230, 255], np.uint8), nblk), 1024)[:L.host_size]
    keep = rng.integers(0, 255, L.host_size, dtype=np.uint8) < level
    bp_np = np.maximum(rng.integers(0, 256, L.host_size, dtype=np.uint8), 1) * keep.astype(np.uint8)
    s_o, cum_o = oracle.lossless_encode(bp_np, Lo)
    out, s_g, cum_g = _gpu_encode(_lib, L, bp_np)
    assert s_g.size + (-s_g.size) % 8 == s_o.size
    assert np.array_equal(s_g, s_o[:s_g.size])
    assert np.array_equal(cum_g, cum_o)
    bp = torch.zeros(L.host_size, dtype=torch.uint8, device="cuda")
    ws = torch.empty(_lib.lib().pa_decode_workspace(C.byref(L)), dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().pa_lossless_decode(P(out), out.numel(), C.byref(L), P(bp), P(ws), ws.numel(), S()))
    assert np.array_equal(host(bp), bp_np)
    assert int(host(ws[:4].view(torch.int32))[0]) == 0
    b, e = [0, 3, 17, 30], [12, 32, 24, 31]
    be = dev(np.array(b + e, np.int32))
    bp2 = torch.zeros(L.host_size, dtype=torch.uint8, device="cuda")
    nread = torch.zeros(1, dtype=torch.int64, device="cuda")
    _lib.check(_lib.lib().pa_lossless_decode_window(P(out), out.numel(), C.byref(L), P(be[0:4]), P(be[4:8]), P(bp2),
                                                     P(nread), 0, P(ws), ws.numel(), S()))
    got = host(bp2)
    ref, nb = oracle.lossless_decode_window(s_o, Lo, b, e, corrected=True)
    assert int(host(nread)[0]) == nb
    for lr in range(4):
        lv = 3 - lr
        lo = L.dev_base[lv] + b[lr] * L.dev_stride[lv]
        hi = L.dev_base[lv] + e[lr] * L.dev_stride[lv]
        assert np.array_equal(got[lo:hi], bp_np[lo:hi])


def test_strategy_parity(pa, oracle):
    _lib, pipe = pa
    rng = np.random.default_rng(5)
    ws = torch.zeros(_lib.lib().pa_strategy_workspace(), dtype=torch.uint8, device="cuda")
    for trial in range(12):
        announce("strategy", trial)
        sizes = rng.integers(2, 6000, 128) if trial % 3 else rng.choice([2, 4], 128)
        cum = np.cumsum(sizes).astype(np.uint64)
        cum_d = dev(cum.view(np.int64))
        st = torch.zeros(9, dtype=torch.int32, device="cuda")
        b_o, e_o = np.zeros(4, np.int32), np.zeros(4, np.int32)
        eb = 1e-6 * (1 + trial)
        for target in (1e-2, 1e-3, 3e-4, 1e-4, 2e-5, 1e-5, 5e-7):
            b_o, e_o, ml_o, _ = oracle.find_strategy(cum, b_o, e_o, eb, target)
            _lib.check(_lib.lib().pa_find_strategy(P(cum_d), P(st[0:4]), P(st[4:8]), P(st[8:9]), eb, target, P(ws),
                                                    ws.numel(), S()))
            h = host(st)
            assert list(h[0:4]) == list(b_o) and list(h[4:8]) == list(e_o) and h[8] == ml_o, (trial, target)
            # the search's best key and arrival count are back at zero for the next call
            assert not host(ws).any(), (trial, target)


CONFIGS = [((33, 47, 29), np.float32, 1e-3), ((16, 16, 16), np.float64, 1e-2), ((1, 1, 5000), np.float64, 1e-4),
           ((1, 70, 90), np.float32, 1e-4), ((64, 64, 64), np.float32, 1e-3), ((40, 50, 60), np.float64, 1e-6),
           ((1, 1, 70001), np.float32, 1e-5), ((1, 300, 257), np.float64, 1e-5)]


@pytest.mark.parametrize("shape,dtype,rel", CONFIGS)
@pytest.mark.parametrize("progressive", [False, True])
def test_end_to_end_stream_parity(pa, oracle, shape, dtype, rel, progressive):
    _lib, pipe = pa
    announce("e2e", shape, dtype.__name__, rel, progressive)
    data = smooth(shape, dtype, seed=7)
    s_o, fl_o = oracle.compress(data, rel, mode=0 if progressive else 1)
    c = pipe.compress(dev(data), rel, rel=True, progressive=progressive)
    s_g = host(c.stream)
    assert s_g.size == s_o.size
    assert np.array_equal(s_g, s_o)
    if not progressive:
        r_o = oracle.decompress(s_o, shape, dtype, rel)
        r_g = host(pipe.decompress(c))
        assert np.array_equal(r_g.view(np.uint8), r_o.view(np.uint8))
    else:
        po = oracle.Progressive(s_o, shape, dtype, rel, corrected=True)
        rv = pipe.Retriever(c)
        for t in (1e-1, 1e-2, 1e-3, 1e-4, rel, rel / 2):
            r_o, ml_o, _ = po.retrieve(t)
            r_g = host(rv.retrieve(t))
            assert np.array_equal(r_g.view(np.uint8), r_o.view(np.uint8)), t


def test_grf_progressive_parity(pa, oracle):
    """The 512^3 float32 Gaussian-random-field workload (rel 1e-4, retrieval
    at 1e-1 .. 1e-4) at 80^3: stream and every round bit-equal to the
    oracle."""
    _lib, pipe = pa
    from prism_amd import fields
    f = fields.grf(80, 80, 80, seed=1, dtype=torch.float32, device="cuda")
    data = host(f)
    announce("grf progressive 80^3 f32 rel 1e-4")
    s_o, _ = oracle.compress(data, 1e-4, mode=0)
    c = pipe.compress(f, 1e-4, rel=True, progressive=True)
    assert np.array_equal(host(c.stream), s_o)
    po = oracle.Progressive(s_o, data.shape, data.dtype, 1e-4, corrected=True)
    rv = pipe.Retriever(c)
    for t in (1e-1, 1e-2, 1e-3, 1e-4):
        r_o, _, _ = po.retrieve(t)
        assert np.array_equal(host(rv.retrieve(t)).view(np.uint8), r_o.view(np.uint8)), t


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_quality_statistics(pa, dtype):
    """Device statistics against numpy: L-inf and its first index exact,
    min / max exact, sums within rounding; ties resolve to the first index."""
    _lib, pipe = pa
    rng = np.random.default_rng(3)
    for n in (1, 257, (1 << 20) + 3):
        announce("quality", dtype.__name__, n)
        a = rng.normal(0, 1, n).astype(dtype)
        b = (a + rng.normal(0, 1e-3, n)).astype(dtype)
        if n > 300:  # two equal largest errors: the first index wins
            b[100] = a[100] + dtype(0.5)
            b[200] = a[200] + dtype(0.5)
        q = pipe.quality(dev(a), dev(b))
        e = np.abs(b.astype(np.float64) - a.astype(np.float64))
        assert q["linf"] == e.max() and q["linf_index"] == int(np.argmax(e))
        assert q["min"] == float(a.min()) and q["max"] == float(a.max())
        sse = float((e * e).sum())
        assert abs(q["sse"] - sse) <= 1e-12 * max(sse, 1e-300)
        rg = float(a.max()) - float(a.min())
        if rg > 0 and sse > 0:
            mse = sse / n
            assert abs(q["psnr_db"] - (20 * np.log10(rg) - 10 * np.log10(mse))) < 1e-9
            assert abs(q["nrmse"] - np.sqrt(mse) / rg) <= 1e-12 * np.sqrt(mse) / rg


HANDOFF_CASES = [((40, 50, 60), np.float64, (1e-2, 1e-3, 1e-4, 1e-5, 1e-6)),
                 ((33, 17, 18), np.float64, (1e-1, 1e-3, 1e-5, 1e-6)),
                 ((65, 47, 34), np.float32, (1e-1, 1e-2, 1e-3, 1e-4)),
                 ((1, 80, 96), np.float64, (1e-2, 1e-4, 1e-6))]  # (nz, ny, nx)


@pytest.mark.parametrize("shape,dtype,targets", HANDOFF_CASES,
                         ids=["x".join(map(str, c[0])) + "-" + np.dtype(c[1]).name for c in HANDOFF_CASES])
def test_code_handoffs_equal_int32(pa, oracle, shape, dtype, targets):
    """The three ways a progressive round's codes reach the reconstruction
    give the same bits as the int32 stage calls: the Retriever's (the
    reconstruction reads the decoded bit-planes when every level's window
    holds at most 8 planes; else an int16 or int32 hand-off) and the int16
    hand-off entries (every level reading planes >= 17).  The first round
    (planes from 0) hands over int32 codes; 2-D fields never read planes."""
    _lib, pipe = pa
    data = smooth(shape, dtype, seed=3)
    c = pipe.compress(dev(data), 1e-6 if dtype == np.float64 else 1e-4, rel=True, progressive=True)
    r = pipe.Retriever(c)
    L = r.L
    dt = _lib.PA_F64 if dtype == np.float64 else _lib.PA_F32
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    flags_seen = []
    st32 = torch.zeros(16, dtype=torch.int32, device="cuda")
    bp32 = torch.empty(L.host_size, dtype=torch.uint8, device="cuda")
    qc32 = torch.empty(L.code_len, dtype=torch.int32, device="cuda")
    qc16 = torch.empty(L.code_len, dtype=torch.int32, device="cuda")
    src16 = torch.zeros(1, dtype=torch.int32, device="cuda")
    dws = torch.empty(_lib.lib().pa_decode_workspace(C.byref(L)), dtype=torch.uint8, device="cuda")
    sws = torch.zeros(_lib.lib().pa_strategy_workspace(), dtype=torch.uint8, device="cuda")
    old = old16 = None
    lib = _lib.lib()
    for k, t in enumerate(targets):
        announce("code hand-off round", shape, k, t)
        got = r.retrieve(t)
        flags_seen.append(int(r.narrow.item()))
        _lib.check(lib.pa_find_strategy(P(r.cum), P(st32[0:4]), P(st32[4:8]), P(st32[8:9]), r.eb_abs, t * r.range,
                                        P(sws), sws.numel(), S()))
        _lib.check(lib.pa_lossless_decode_window(P(c.stream), c.nbytes, C.byref(L), P(st32[0:4]), P(st32[4:8]),
                                                 P(bp32), None, int(k > 0), P(dws), dws.numel(), S()))
        _lib.check(lib.pa_inverse_bitshuffle_window(P(bp32), C.byref(L), _lib.PA_MODE_NB, P(st32[0:4]), P(st32[4:8]),
                                                    P(qc32), S()))
        _lib.check(lib.pa_inverse_bitshuffle_window_narrow(P(bp32), C.byref(L), P(st32[0:4]), P(st32[4:8]), P(qc16),
                                                           P(src16), S()))
        new = torch.empty(shape, dtype=tdt, device="cuda")
        new16 = torch.empty_like(new)
        if old is None:
            _lib.check(lib.pa_spline_reconstruct(P(qc32), P(r.anchors), dt, C.byref(L), r.eb_abs, r.rel_eb,
                                                 P(r.flags), P(st32[8:9]), P(new), S()))
            _lib.check(lib.pa_spline_reconstruct_narrow(P(qc16), P(src16), P(r.anchors), dt, C.byref(L), r.eb_abs,
                                                        r.rel_eb, P(r.flags), P(st32[8:9]), P(new16), S()))
        else:
            _lib.check(lib.pa_spline_delta_reconstruct(P(qc32), dt, C.byref(L), r.eb_abs, r.rel_eb,
                                                       P(r.flags), P(st32[8:9]), P(old), P(new), S()))
            _lib.check(lib.pa_spline_delta_reconstruct_narrow(P(qc16), P(src16), dt, C.byref(L), r.eb_abs, r.rel_eb,
                                                              P(r.flags), P(st32[8:9]), P(old16), P(new16), S()))
        old, old16 = new, new16
        assert torch.equal(got.view(torch.uint8), new.view(torch.uint8)), k
        assert torch.equal(new16.view(torch.uint8), new.view(torch.uint8)), k
        assert torch.equal(r.state[0:9], st32[0:9])
    assert flags_seen[0] == 0, flags_seen
    if shape[0] > 1:  # nz > 1
        assert 2 in flags_seen[1:], flags_seen
    else:
        assert 2 not in flags_seen, flags_seen


# (begin[4], end[4]) per level lr = 3 - l: the 8 top planes, the 8 last,
# mixed widths with an empty level, every level empty, odd offsets
PLANE_WINDOWS = [([0, 0, 0, 0], [8, 8, 8, 8]), ([24, 25, 26, 27], [32, 32, 32, 32]), ([31, 5, 9, 0], [32, 5, 17, 8]),
                 ([3, 3, 3, 3], [3, 3, 3, 3]), ([12, 0, 30, 17], [20, 3, 32, 25])]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(33, 47, 29), (16, 16, 16), (35, 18, 20), (17, 17, 17), (40, 64, 64), (6, 1, 70),
                                   (1, 40, 70)])
def test_plane_reads_equal_int32_handoff(pa, shape, dtype):
    """The reconstruction that reads its codes from the bit-planes equals the
    int32 hand-off (windowed inverse shuffle, then the reconstruction) on
    random plane bytes, under chosen windows of at most 8 planes per level
    (the first and the last planes, an empty level, all levels empty), for
    maxlevel 0 and 3, in the anchored and the delta form.  A 2-D field is
    never deferred: its reconstruction reads the int32 or int16 hand-off."""
    _lib, pipe = pa
    nz, ny, nx = shape
    L = _lib.layout(nx, ny, nz)
    dt = _lib.PA_F32 if dtype == np.float32 else _lib.PA_F64
    lib = _lib.lib()
    rng = np.random.default_rng(nx * 7 + ny)
    bp = dev(rng.integers(0, 256, L.host_size, dtype=np.uint8))
    an = dev(rng.normal(0, 1, L.n4).astype(dtype))
    old = dev(rng.normal(0, 1, shape).astype(dtype))
    fl = dev(np.array([1, 0, 1, 0, 0, 1, 0, 1], np.uint8))
    eb, rel = 1e-3, 1e-4
    qc32 = torch.zeros(L.code_len, dtype=torch.int32, device="cuda")
    qcd = torch.zeros(L.code_len, dtype=torch.int32, device="cuda")
    src = torch.zeros(1, dtype=torch.int32, device="cuda")
    for b, e in PLANE_WINDOWS:
        for maxlevel in (0, 3):
            announce("plane reads", shape, dtype.__name__, b, e, maxlevel)
            st = dev(np.array(b + e + [maxlevel], np.int32))
            B, E, ML = st[0:4], st[4:8], st[8:9]
            _lib.check(lib.pa_inverse_bitshuffle_window(P(bp), C.byref(L), _lib.PA_MODE_NB, P(B), P(E), P(qc32), S()))
            _lib.check(lib.pa_inverse_bitshuffle_window_deferred(P(bp), C.byref(L), P(B), P(E), P(qcd), P(src), S()))
            assert int(host(src)[0]) == (2 if nz > 1 else int(min(b[lr] >= 17 or b[lr] >= e[lr]
                                                                   for lr in range(4))))
            for form in ("anchored", "delta"):
                want = torch.empty(shape, dtype=old.dtype, device="cuda")
                got = torch.empty_like(want)
                if form == "anchored":
                    _lib.check(lib.pa_spline_reconstruct(P(qc32), P(an), dt, C.byref(L), eb, rel, P(fl), P(ML),
                                                         P(want), S()))
                    _lib.check(lib.pa_spline_reconstruct_planes(P(qcd), P(src), P(bp), P(B), P(E), P(an), None, dt,
                                                                C.byref(L), eb, rel, P(fl), P(ML), P(got), S()))
                else:
                    _lib.check(lib.pa_spline_delta_reconstruct(P(qc32), dt, C.byref(L), eb, rel, P(fl), P(ML),
                                                               P(old), P(want), S()))
                    _lib.check(lib.pa_spline_reconstruct_planes(P(qcd), P(src), P(bp), P(B), P(E), None, P(old), dt,
                                                                C.byref(L), eb, rel, P(fl), P(ML), P(got), S()))
                assert torch.equal(got.view(torch.uint8), want.view(torch.uint8)), (b, e, maxlevel, form)
    # exactly one of anchors and old
    assert lib.pa_spline_reconstruct_planes(P(qcd), P(src), P(bp), P(B), P(E), P(an), P(old), dt, C.byref(L), eb,
                                            rel, P(fl), P(ML), P(old), S()) == _lib.PA_ERR_ARG


# ---- the closed-form pins of the tuner's sample geometry and of the
# partial-tile rule selection, run through the HIP kernels ----

from closed_forms import (RULE_COEFS, RULE_SHAPES, TUNER_SHAPES, codes_by_point, cubic_axes_field,  # noqa: E402
                          expected_codes, noise_outside_windows, tuner_blocks)


def _gpu_scores(_lib, d, L, dt, both=True):
    fl = torch.empty(8, dtype=torch.uint8, device="cuda")
    sc = torch.empty(15, dtype=torch.float64, device="cuda")
    ws = torch.empty(max(8, _lib.lib().pa_autotune_workspace(C.byref(L))), dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().pa_autotune(P(d), dt, C.byref(L), P(sc), P(fl), P(ws), ws.numel(), S()))
    out = [host(sc)]
    if both:
        mm = torch.empty(4, dtype=torch.float64, device="cuda")
        sc2 = torch.empty(15, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().pa_extrema_autotune(P(d), dt, C.byref(L), P(mm), 1e-4, None, P(sc2), P(fl), P(ws),
                                                  ws.numel(), S()))
        out.append(host(sc2))
    return out


@pytest.mark.parametrize("shape_xyz", TUNER_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_tuner_sample_windows(pa, oracle, shape_xyz):
    """Both tuner entries read exactly the hand-derived sample windows:
    zero scores on noise outside them, positive scores for an impulse at
    either corner of every window, and scores bit-equal to the oracle's."""
    _lib, pipe = pa
    nx, ny, nz = shape_xyz
    L = _lib.layout(nx, ny, nz)
    announce("tuner windows", shape_xyz)
    f = noise_outside_windows(shape_xyz, np.float32)
    d = dev(f)
    sc_o, _, _ = oracle.autotune(f)
    for sc in _gpu_scores(_lib, d, L, _lib.PA_F32):
        assert not sc.any() and np.array_equal(sc.view(np.uint64), sc_o.view(np.uint64))
    d.zero_()
    for ox, oy, oz in tuner_blocks(shape_xyz):
        far = [o + min(16, 8 * ((n - 1 - o) // 8)) for o, n in ((ox, nx), (oy, ny), (oz, nz))]
        for px, py, pz in ((ox, oy, oz), tuple(far)):
            announce("tuner impulse", shape_xyz, (px, py, pz))
            d[pz, py, px] = 1.0
            for sc in _gpu_scores(_lib, d, L, _lib.PA_F32, both=(px, py, pz) == (ox, oy, oz)):
                assert (sc > 0).all(), ((px, py, pz), sc)
            d[pz, py, px] = 0.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", RULE_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_partial_tile_rules(pa, oracle, dtype, shape):
    """Codes of the one-axis cubic fields equal the closed-form rule
    remainders (and the oracle's), and both reconstructions are exact."""
    _lib, pipe = pa
    nz, ny, nx = shape
    L = _lib.layout(nx, ny, nz)
    dt = _lib.PA_F32 if dtype == np.float32 else _lib.PA_F64
    for coef in RULE_COEFS:
        data = cubic_axes_field(shape, coef, dtype)
        d = dev(data)
        for flags in FLAGSETS + [[1] * 8]:
            announce("partial-tile rules", shape, dtype.__name__, coef, flags)
            fl = dev(np.array(flags, np.uint8))
            qc = torch.empty(L.code_len, dtype=torch.int32, device="cuda")
            an = torch.empty(L.n4, dtype=d.dtype, device="cuda")
            _lib.check(_lib.lib().pa_spline_construct(P(d), dt, C.byref(L), 0.25, 0.0, P(fl), P(qc), P(an), S()))
            got = host(qc)
            assert np.array_equal(codes_by_point(L, got, shape), expected_codes(shape, coef, flags))
            qc_o, _ = oracle.spline_construct(data, 0.25, 0.0, flags)
            assert np.array_equal(got, qc_o)
            out = torch.empty(shape, dtype=d.dtype, device="cuda")
            _lib.check(_lib.lib().pa_spline_reconstruct(P(qc), P(an), dt, C.byref(L), 0.25, 0.0, P(fl), None,
                                                         P(out), S()))
            assert np.array_equal(host(out), data)
