"""Pins of the oracle's stream geometry.

* sizes of eight shapes against the size table derived from the reference's
  formulas (tests/golden/layout_sizes.json);
* the level-major code index (InterpolationPredictor.cu:2070-2083) against
  its definition: anchors first in raster order, then each level's points
  (coarsest first) in raster order, every level region padded to a multiple
  of 8 codes; checked by brute force on small fields.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sizes_match_golden(oracle):
    with open(os.path.join(HERE, "golden", "layout_sizes.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        L = oracle.layout(*c["shape"])
        assert L.n4 == c["n4"], c
        assert list(L.host_stride) == c["e"], c
        assert list(L.dev_stride) == c["s"], c
        assert L.nchunks == c["chunks"], c
        for key in ("trailer_f32", "trailer_f64"):
            if key in c:
                s = 4 if key.endswith("32") else 8
                assert L.trailer_fixed + L.n4 * s == c[key]


def _level_of(gx, gy, gz):
    lvl = 0
    while lvl < 4 and gx % 2 == 0 and gy % 2 == 0 and gz % 2 == 0:
        gx, gy, gz, lvl = gx >> 1, gy >> 1, gz >> 1, lvl + 1
    return lvl


@pytest.mark.parametrize("shape", [(1, 1, 1), (17, 1, 1), (40, 1, 1), (33, 20, 1), (17, 17, 17),
                                   (33, 47, 29), (16, 16, 16), (5, 70, 3)])
def test_level_major_index_is_rank_within_level(oracle, shape):
    nx, ny, nz = shape
    L = oracle.layout(nx, ny, nz)
    by_level = {l: [] for l in range(5)}
    for gz in range(nz):
        for gy in range(ny):
            for gx in range(nx):
                by_level[_level_of(gx, gy, gz)].append(oracle.lm_index(L, gx, gy, gz))
    # anchors: raster order of the anchor grid, 0..n4-1
    assert by_level[4] == list(range(L.n4))
    # level regions: coarsest first, each padded to a multiple of 8
    start = L.n4
    for lvl in (3, 2, 1, 0):
        idx = by_level[lvl]
        assert idx == list(range(start, start + len(idx))), (shape, lvl)
        assert L.prefix[lvl] == start
        start += len(idx)
        if lvl > 0:
            start += (-len(idx)) % 8
    assert L.real_end == start
    # total count and device segment lengths (pads + the n4 phantom codes)
    assert sum(len(v) for v in by_level.values()) == nx * ny * nz
    assert L.dev_len[0] == nx * ny * nz - L.cnt[0] + L.n4 + (-(nx * ny * nz - L.cnt[0] + L.n4)) % 8
    for l in range(4):
        assert L.dev_len[l] % 8 == 0


def test_host_strides_exceed_device_strides(oracle):
    """io.hpp:121 always pads by 1..8 codes where bitplane.cu:31 pads by 0..7,
    so e_l = s_l or s_l + 8; the host buffer therefore covers the device one."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        nx, ny, nz = (int(v) for v in rng.integers(1, 300, 3))
        L = oracle.layout(nx, ny, nz)
        for l in range(4):
            assert L.host_stride[l] - L.dev_stride[l] in (0, 8)
            assert L.host_base[l] >= L.dev_base[l]
        assert L.host_size >= L.dev_size
