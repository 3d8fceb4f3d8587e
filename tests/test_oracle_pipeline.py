"""End-to-end pins of the oracle's stream path on the 64^3 float32 field of
8 random-phase sinusoids + 4 Gaussian blobs + N(0, 1e-3) noise (seed 0),
relative bound 1e-3: the reference's CPU-sized plain case, plus a
progressive run of the same field.

What fixes the bytes here without the oracle's own formulas:
* stream layout (zbpc.cu:476-520, compressor.cu:139-151): the header word is
  the bit-plane buffer size, the size table has one entry per chunk, the
  trailer holds the flags, max - min of the field, the size table by
  (level, plane) and the raw values at the 16-multiples;
* the decompressed field is within the bound at every point;
* progressive rounds meet each target, and retrieving down to the
  compression bound reads every payload byte exactly once.
"""
import numpy as np
import pytest
import torch

from prism_amd import fields


@pytest.fixture(scope="module")
def cfg1():
    return np.ascontiguousarray(fields.smooth_mix(64, 64, 64, seed=0, dtype=torch.float32, device="cpu").numpy())


def _split(stream, L, s):
    C = int(L.nchunks)
    insize = int(stream[:8].view(np.uint64)[0])
    csize = stream[8:8 + 2 * C].view(np.uint16).astype(np.int64)
    tail = stream[len(stream) - (int(L.trailer_fixed) + int(L.n4) * s):]
    flags = tail[:8]
    rng = float(tail[8:16].view(np.float64)[0])
    cum = tail[16:16 + 1024].view(np.uint64).reshape(4, 32)
    anchors = tail[1040:]
    return insize, csize, flags, rng, cum, anchors


@pytest.mark.parametrize("mode", [1, 0], ids=["sm", "nb"])
def test_stream_layout(oracle, cfg1, mode):
    stream, flags = oracle.compress(cfg1, 1e-3, mode=mode)
    L = oracle.layout(64, 64, 64)
    insize, csize, fl, rng, cum, anchors = _split(stream, L, 4)
    assert insize == L.host_size
    assert np.array_equal(fl, flags)
    assert rng == float(cfg1.max()) - float(cfg1.min())
    assert np.array_equal(anchors.view(np.float32), cfg1[::16, ::16, ::16].ravel())
    payload = int(csize.sum())
    assert int(cum[3, 31]) == payload  # last (level, plane) closes the payload
    assert np.all(np.diff(cum.ravel().astype(np.int64)) >= 0)
    body = 8 + 2 * int(L.nchunks) + payload
    assert len(stream) == (body + 7) // 8 * 8 + int(L.trailer_fixed) + int(L.n4) * 4
    assert not stream[body:(body + 7) // 8 * 8].any()


def test_plain_roundtrip_meets_bound(oracle, cfg1):
    stream, _ = oracle.compress(cfg1, 1e-3, mode=1)
    out = oracle.decompress(stream, cfg1.shape, cfg1.dtype, 1e-3)
    rg = float(cfg1.max()) - float(cfg1.min())
    err = np.abs(out.astype(np.float64) - cfg1.astype(np.float64)).max()
    assert err <= 1e-3 * rg * (1 + 1e-5)
    assert len(stream) < cfg1.nbytes / 3


def test_progressive_rounds(oracle, cfg1):
    """Corrected selection: every round meets its target and the payload is
    read at least once.  The reference's selection (host chunks
    [begin*cb, end*cb) per level) takes disjoint windows, so down to the
    compression bound it reads every payload byte exactly once."""
    stream, _ = oracle.compress(cfg1, 1e-3, mode=0)
    L = oracle.layout(64, 64, 64)
    *_, cum, _ = _split(stream, L, 4)
    payload = int(cum[3, 31])
    rg = float(cfg1.max()) - float(cfg1.min())
    for corrected in (True, False):
        p = oracle.Progressive(stream, cfg1.shape, cfg1.dtype, 1e-3, corrected=corrected)
        read = 0
        for t in (1e-1, 1e-2, 1e-3):
            out, _, nb = p.retrieve(t)
            read += nb
            err = np.abs(out.astype(np.float64) - cfg1.astype(np.float64)).max()
            if corrected:
                assert err <= t * rg * (1 + 1e-5), (t, err / (t * rg))
        assert p.end.tolist() == [32] * 4
        if corrected:
            assert read >= payload
        else:
            assert read == payload
