"""Pins of the oracle's chunk codec and lossless stream.

The expected bytes of the hand example below follow the format stated in
RZE.hpp:10-86 step by step (zero elimination with an LSB-first bitmap, three
rounds of repetition elimination over 2048 -> 256 -> 32 -> 4 bitmap bytes,
the 4-byte top bitmap, the 2-byte size footer).
"""
import numpy as np
import pytest

CS = 16384


def test_hand_example(oracle):
    chunk = np.zeros(16, np.uint8)
    chunk[3] = 7
    chunk[10] = 7
    enc = oracle.rze_encode_chunk(chunk)
    expected = [7, 7,            # non-zero bytes
                0x08, 0x04, 0x00,  # B0 bytes differing from their predecessor: B0[0..2]
                0x07, 0x00,        # B1 = [0x07, 0, ...]: bytes 0, 1 differ
                0x03, 0x00,        # B2 = [0x03, 0, ...]
                0x03, 0x00, 0x00, 0x00,  # B3 (4 bytes)
                16, 0]             # original size, little endian
    assert enc is not None and list(enc) == expected
    assert list(oracle.rze_decode_chunk(enc)) == list(chunk)


def test_all_zero_chunk_is_footer_only(oracle):
    for n in (8, 4096, CS):
        enc = oracle.rze_encode_chunk(np.zeros(n, np.uint8))
        assert list(enc) == [n & 0xFF, n >> 8]
        assert not oracle.rze_decode_chunk(enc).any()


def test_incompressible_chunk_is_raw(oracle):
    rng = np.random.default_rng(1)
    data = rng.integers(1, 256, CS).astype(np.uint8)
    assert oracle.rze_encode_chunk(data) is None


def test_overflow_limits(oracle):
    """RZE.hpp:34: the zero-eliminated data alone must stay below CS-2 bytes."""
    rng = np.random.default_rng(2)
    # 16381 non-zero bytes + bitmaps can never fit: raw
    data = np.zeros(CS, np.uint8)
    data[:16381] = rng.integers(1, 256, 16381)
    assert oracle.rze_encode_chunk(data) is None
    # few non-zero bytes always fit
    data = np.zeros(CS, np.uint8)
    data[::97] = 5
    enc = oracle.rze_encode_chunk(data)
    assert enc is not None and len(enc) < CS
    assert np.array_equal(oracle.rze_decode_chunk(enc), data)


@pytest.mark.parametrize("density", [0.0, 0.001, 0.05, 0.3, 0.7])
@pytest.mark.parametrize("size", [8, 64, 1000, 8192, CS])
def test_roundtrip(oracle, density, size):
    rng = np.random.default_rng(int(density * 1000) + size)
    for _ in range(3):
        data = np.where(rng.random(size) < density, rng.integers(1, 256, size), 0).astype(np.uint8)
        enc = oracle.rze_encode_chunk(data)
        if enc is None:
            continue
        assert len(enc) >= 2
        assert int(enc[-2]) | (int(enc[-1]) << 8) == size
        assert np.array_equal(oracle.rze_decode_chunk(enc), data)


@pytest.mark.parametrize("shape", [(64, 64, 64), (33, 47, 29), (300, 1, 1), (70, 50, 1)])
def test_stream_roundtrip_and_table(oracle, shape):
    L = oracle.layout(*shape)
    rng = np.random.default_rng(7)
    bp = np.where(rng.random(L.host_size) < 0.05, rng.integers(1, 256, L.host_size), 0).astype(np.uint8)
    stream, cum = oracle.lossless_encode(bp, L)
    assert stream.size % 8 == 0
    assert int(np.frombuffer(stream[:8].tobytes(), np.uint64)[0]) == L.host_size
    sizes = np.frombuffer(stream[8:8 + 2 * L.nchunks].tobytes(), np.uint16).astype(np.int64)
    assert (sizes >= 2).all()
    payload = int(sizes.sum())
    assert stream.size == 8 + 2 * L.nchunks + payload + (-(8 + 2 * L.nchunks + payload)) % 8
    # cum[lr][b]: inclusive payload bytes through the last chunk of (lr, b)
    ends = np.cumsum(sizes)
    k = 0
    for lr in range(4):
        cb = L.chunks_per_plane[3 - lr]
        for b in range(32):
            k += cb
            assert cum[lr * 32 + b] == ends[k - 1]
    assert np.array_equal(oracle.lossless_decode(stream, L), bp)


def test_window_decode_selects_planes(oracle):
    L = oracle.layout(40, 40, 40)
    rng = np.random.default_rng(8)
    bp = np.where(rng.random(L.host_size) < 0.2, rng.integers(1, 256, L.host_size), 0).astype(np.uint8)
    stream, cum = oracle.lossless_encode(bp, L)
    begin, end = [2, 5, 0, 31], [9, 5, 32, 32]
    got, nbytes = oracle.lossless_decode_window(stream, L, begin, end, corrected=False)
    sizes = np.frombuffer(stream[8:8 + 2 * L.nchunks].tobytes(), np.uint16).astype(np.int64)
    want_bytes = 0
    k = 0
    for lr in range(4):
        l = 3 - lr
        e, cb = L.host_stride[l], L.chunks_per_plane[l]
        for b in range(32):
            seg = slice(L.host_base[l] + b * e, L.host_base[l] + (b + 1) * e)
            if begin[lr] <= b < end[lr]:
                assert np.array_equal(got[seg], bp[seg])
                want_bytes += int(sizes[k:k + cb].sum())
            else:
                assert not got[seg].any()
            k += cb
    assert nbytes == want_bytes
    # the covering selection decodes every device byte of the requested planes
    got2, _ = oracle.lossless_decode_window(stream, L, begin, end, corrected=True)
    for lr in range(4):
        l = 3 - lr
        if begin[lr] < end[lr]:
            lo = L.dev_base[l] + begin[lr] * L.dev_stride[l]
            hi = L.dev_base[l] + end[lr] * L.dev_stride[l]
            assert np.array_equal(got2[lo:hi], bp[lo:hi])
