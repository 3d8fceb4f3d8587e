"""Pins of the oracle's bit-plane shuffle.

The oracle states the shuffle as a closed form (byte (l,k,g) collects bit
31-k of the eight codes of group g, code j at bit 7-j).  Here that closed
form is checked against an emulation, written from the reference's device
code, of the warp network that actually produces the bytes:
bitplane.cu:90-109 (three __shfl_xor_sync + __byte_perm steps and three
delta swaps) and the byte placement at 120-132; and of the inverse network
at 292-318.
"""
import numpy as np
import pytest

M32 = 0xFFFFFFFF


def byte_perm(x, y, s):
    """CUDA __byte_perm: result byte n = byte (s >> 4n) & 7 of the 8 bytes {y:x}."""
    src = [(x >> (8 * i)) & 0xFF for i in range(4)] + [(y >> (8 * i)) & 0xFF for i in range(4)]
    r = 0
    for n in range(4):
        r |= src[(s >> (4 * n)) & 7] << (8 * n)
    return r


def warp_network(a):
    """bitplane.cu:90-109 for the 32 lanes of one warp."""
    a = list(a)
    q = [a[l ^ 4] for l in range(32)]
    a = [byte_perm(a[l], q[l], 0x3276) if (l & 4) == 0 else byte_perm(a[l], q[l], 0x5410) for l in range(32)]
    q = [a[l ^ 2] for l in range(32)]
    a = [byte_perm(a[l], q[l], 0x3715) if (l & 2) == 0 else byte_perm(a[l], q[l], 0x6240) for l in range(32)]
    q = [a[l ^ 1] for l in range(32)]
    m = 0x0F0F0F0F
    a = [((a[l] & ~m) | ((q[l] >> 4) & m)) & M32 if (l & 1) == 0 else (((q[l] << 4) & ~m) | (a[l] & m)) & M32
         for l in range(32)]
    out = []
    for v in a:
        t = (v ^ (v >> 3)) & 0x0A0A0A0A
        v = v ^ t ^ (t << 3)
        t = (v ^ (v >> 6)) & 0x00CC00CC
        v = v ^ t ^ (t << 6)
        t = (v ^ (v >> 12)) & 0x0000F0F0
        v = v ^ t ^ (t << 12)
        out.append(v & M32)
    return out


def inverse_warp_network(a):
    """bitplane.cu:298-318."""
    b = []
    for v in a:
        t = (v ^ (v >> 12)) & 0x0000F0F0
        v = v ^ t ^ (t << 12)
        t = (v ^ (v >> 6)) & 0x00CC00CC
        v = v ^ t ^ (t << 6)
        t = (v ^ (v >> 3)) & 0x0A0A0A0A
        v = v ^ t ^ (t << 3)
        b.append(v & M32)
    a = b
    q = [a[l ^ 4] for l in range(32)]
    a = [byte_perm(a[l], q[l], 0x3276) if (l & 4) == 0 else byte_perm(a[l], q[l], 0x5410) for l in range(32)]
    q = [a[l ^ 2] for l in range(32)]
    a = [byte_perm(a[l], q[l], 0x3715) if (l & 2) == 0 else byte_perm(a[l], q[l], 0x6240) for l in range(32)]
    q = [a[l ^ 1] for l in range(32)]
    m = 0x0F0F0F0F
    return [((a[l] & ~m) | ((q[l] >> 4) & m)) & M32 if (l & 1) == 0 else (((q[l] << 4) & ~m) | (a[l] & m)) & M32
            for l in range(32)]


def closed_form_bytes(w8):
    """32 plane bytes of one 8-word group: plane k, word j -> bit 7-j."""
    out = []
    for k in range(32):
        b = 0
        for j in range(8):
            b |= ((w8[j] >> (31 - k)) & 1) << (7 - j)
        out.append(b)
    return out


def test_closed_form_matches_reference_network():
    rng = np.random.default_rng(11)
    for trial in range(300):
        if trial % 3 == 0:
            w = [int(v) for v in rng.integers(0, 2**32, 32, dtype=np.uint64)]
        elif trial % 3 == 1:
            w = [int(v) & 0xFF for v in rng.integers(0, 2**32, 32, dtype=np.uint64)]
        else:  # single set bits
            w = [1 << int(rng.integers(0, 32)) if rng.random() < 0.3 else 0 for _ in range(32)]
        a = warp_network(w)
        for grp in range(4):
            # element e = 8*grp + j holds planes 4j..4j+3 (bitplane.cu:127-132)
            got = [0] * 32
            for j in range(8):
                v = a[8 * grp + j]
                got[4 * j + 0] = (v >> 24) & 0xFF
                got[4 * j + 1] = (v >> 8) & 0xFF
                got[4 * j + 2] = (v >> 16) & 0xFF
                got[4 * j + 3] = v & 0xFF
            assert got == closed_form_bytes(w[8 * grp: 8 * grp + 8])


def test_inverse_network_inverts_closed_form():
    rng = np.random.default_rng(12)
    for _ in range(200):
        w = [int(v) for v in rng.integers(0, 2**32, 32, dtype=np.uint64)]
        lanes = []
        for grp in range(4):
            planes = closed_form_bytes(w[8 * grp: 8 * grp + 8])
            for j in range(8):  # byte assembly of bitplane.cu:292-295
                lanes.append((planes[4 * j] << 24) | (planes[4 * j + 1] << 8) | (planes[4 * j + 2] << 16)
                             | planes[4 * j + 3])
        assert inverse_warp_network(lanes) == w


def _nb(v):
    return ((v + 0xAAAAAAAA) ^ 0xAAAAAAAA) & M32


def test_transform_properties():
    vals = list(range(-70000, 70000, 7)) + [2**29 - 1, -(2**29)]
    for v in vals:
        u = v & M32
        nb = _nb(u)
        assert (((nb ^ 0xAAAAAAAA) - 0xAAAAAAAA) & M32) == u  # NB inverse
        assert nb >> 31 == 0  # plane 0 is zero for |v| < 2^29
        # negabinary digits: value = sum nb_i (-2)^i
        assert sum(((nb >> i) & 1) * (-2) ** i for i in range(32)) == v
        sign = M32 if v < 0 else 0
        sm = (u & 0x7FFFFFFF) ^ sign
        s2 = M32 if sm >> 31 else 0
        assert ((sm & 0x7FFFFFFF) ^ s2) == u  # SM is an involution


@pytest.mark.parametrize("shape", [(64, 64, 64), (33, 47, 29), (17, 1, 1), (100, 30, 1), (16, 16, 16)])
@pytest.mark.parametrize("mode", [0, 1])
def test_roundtrip_and_zero_tail(oracle, shape, mode):
    L = oracle.layout(*shape)
    rng = np.random.default_rng(5)
    qc = np.zeros(L.code_len, np.int32)
    qc[L.n4:L.real_end] = rng.integers(-3000, 3000, L.real_end - L.n4)
    # pads between level regions are not codes
    for lvl, cnt in ((3, L.cnt[2] - L.n4), (2, L.cnt[1] - L.cnt[2]), (1, L.cnt[0] - L.cnt[1])):
        qc[L.prefix[lvl] + cnt: L.prefix[lvl - 1]] = 0
    bp = oracle.bitshuffle(qc, L, mode)
    assert bp.size == L.host_size
    assert not bp[L.dev_size:].any()  # bytes between device and host layout stay zero
    back = oracle.inverse_bitshuffle(bp, L, mode)
    assert np.array_equal(back[L.n4:L.real_end], qc[L.n4:L.real_end])


def test_single_code_placement(oracle):
    L = oracle.layout(40, 40, 40)
    for lvl in range(4):
        for r in (0, 5, 13):
            qc = np.zeros(L.code_len, np.int32)
            start = L.n4 if lvl == 3 else L.prefix[lvl]
            qc[start + r] = -1  # NB(-1) = 0b11, SM(-1) = 0x80000000
            bp = oracle.bitshuffle(qc, L, 0)
            nz = np.flatnonzero(bp)
            g, j = r >> 3, r & 7
            s, a = L.dev_stride[lvl], L.dev_base[lvl]
            assert list(nz) == [a + 30 * s + g, a + 31 * s + g]
            assert bp[a + 31 * s + g] == 1 << (7 - j)
            bp = oracle.bitshuffle(qc, L, 1)
            assert list(np.flatnonzero(bp)) == [a + g]


def test_window_inverse_is_linear(oracle):
    """Digits of disjoint plane windows add up to the code (negabinary)."""
    L = oracle.layout(33, 33, 33)
    rng = np.random.default_rng(9)
    qc = np.zeros(L.code_len, np.int32)
    qc[L.n4:L.real_end] = rng.integers(-10**6, 10**6, L.real_end - L.n4)
    for lvl, cnt in ((3, L.cnt[2] - L.n4), (2, L.cnt[1] - L.cnt[2]), (1, L.cnt[0] - L.cnt[1])):
        qc[L.prefix[lvl] + cnt: L.prefix[lvl - 1]] = 0
    bp = oracle.bitshuffle(qc, L, 0)
    cuts = [0, 7, 13, 20, 26, 32]
    total = np.zeros(L.code_len, np.int64)
    for a, b in zip(cuts[:-1], cuts[1:]):
        part = oracle.inverse_bitshuffle(bp, L, 0, [a] * 4, [b] * 4)
        total += part
    wrapped = ((total[L.n4:L.real_end] + 2**31) % 2**32) - 2**31
    assert np.array_equal(wrapped, qc[L.n4:L.real_end])
