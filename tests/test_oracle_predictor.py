"""Pins of the oracle's multi-level spline predictor and auto-tuner.

* linear fields on 16k+1 grids are predicted exactly by every rule the
  reference uses (cubic, both quadratics, linear, the face fallback), so all
  codes are zero and the reconstruction is exact, for every flag setting;
* impulse responses: with all codes zero except one, reconstruction is pure
  interpolation, and the values at chosen points are derived by hand below
  from the stencils of InterpolationPredictor.cu:1300-1326 (1-D),
  1612-1667 (face) and 1695-1721 (cube);
* the error bound |x - x^| <= eb holds (up to one rounding of T);
* the auto-tuner picks the interpolation order that a counting argument
  predicts for fields that vary along one axis only.
"""
import itertools

import numpy as np
import pytest

ALL_FLAGS = [np.array([md] * 4 + [rv] * 4, np.uint8) for md, rv in itertools.product((0, 1), (0, 1))]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("flags", ALL_FLAGS, ids=lambda f: f"md{f[0]}rev{f[4]}")
def test_linear_field_is_exact(oracle, dtype, flags):
    nz, ny, nx = 33, 17, 49
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    data = (3 * x - 2 * y + 5 * z + 7).astype(dtype)
    qc, anchors = oracle.spline_construct(data, 0.25, 0.0, flags)
    assert not qc.any()
    rec = oracle.spline_reconstruct(qc, anchors, data.shape, 0.25, 0.0, flags)
    assert np.array_equal(rec, data)


def _spike(oracle, shape, point, code, flags, dtype=np.float64):
    """Reconstruction from zero anchors and one non-zero code; eb = 0.5 makes
    ebx2 = 1, so the spiked point reconstructs to pred + code."""
    nz, ny, nx = shape
    L = oracle.layout(nx, ny, nz)
    qc = np.zeros(L.code_len, np.int32)
    qc[oracle.lm_index(L, *point)] = code
    anchors = np.zeros(L.n4, dtype)
    return oracle.spline_reconstruct(qc, anchors, shape, 0.5, 0.0, flags)


@pytest.mark.parametrize("flags", ALL_FLAGS, ids=lambda f: f"md{f[0]}rev{f[4]}")
def test_impulse_response_1d(oracle, flags):
    """Code 1024 at (16,16,8), a unit-8 point of tile (1,1,0) of a 49^3 grid.
    Column x=16,y=16 is predicted along z at every level for every flag
    setting (z-pass, or the z-line of the multi-dimensional rule):
      unit 4: z=4  right quad (3*0 + 6*1024 - 0)/8 = 768
              z=12 left quad (-0 + 6*1024 + 3*0)/8 = 768
      unit 2: z=2  right quad (3*0 + 6*768 - 1024)/8 = 448
              z=6  cubic (-0 + 9*768 + 9*1024 - 768)/16 = 960
              z=10 cubic (-768 + 9*1024 + 9*768 - 0)/16 = 960
              z=14 left quad (-1024 + 6*768 + 3*0)/8 = 448
      unit 1: z=1  right quad (3*0 + 6*448 - 768)/8 = 240
              z=3  cubic (-0 + 9*448 + 9*768 - 960)/16 = 624
              z=5  cubic (-448 + 9*768 + 9*960 - 1024)/16 = 880
    """
    rec = _spike(oracle, (49, 49, 49), (16, 16, 8), 1024, flags)
    col = rec[:, 16, 16]
    want = {0: 0, 8: 1024, 16: 0, 4: 768, 12: 768, 2: 448, 6: 960, 10: 960, 14: 448, 1: 240, 3: 624, 5: 880}
    for z, v in want.items():
        assert col[z] == v, (z, col[z], v)


def test_impulse_response_face_and_cube(oracle):
    """Multi-dimensional rule at unit 2 in the interior tile (1,1,1) of 49^3.
    Code 1024 at tile-local Q=(8,6,8), a y-line point; everything else is 0.
      face P=(8,6,6), directions z and y, both cubic:
          (cubic_z(0,0,1024,0) + cubic_y(0,0,0,0))/2 = (576 + 0)/2 = 288
      face (6,6,8), directions y and x, both cubic:
          (cubic_y(0,0,0,0) + cubic_x(0,0,1024,0))/2 = 288
      cube C=(6,6,6), all three cubic:
          x taps (0,4,8,12) -> (0,0,288,0): 162;  y taps -> all 0: 0;
          z taps (6,6,{0,4,8,12}) -> (0,0,288,0): 162;  (162 + 0 + 162)/3 = 108
    """
    flags = np.array([0, 1, 0, 0, 0, 0, 0, 0], np.uint8)  # unit 2 uses the multi-dimensional rule
    o = 16
    rec = _spike(oracle, (49, 49, 49), (o + 8, o + 6, o + 8), 1024, flags)
    assert rec[o + 8, o + 6, o + 8] == 1024
    assert rec[o + 6, o + 6, o + 8] == 288  # P (x=8,y=6,z=6): rec[z, y, x]
    assert rec[o + 8, o + 6, o + 6] == 288
    assert rec[o + 6, o + 6, o + 6] == 108


def _smooth_field(shape, dtype, seed):
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = np.sin(x / 6.0 + 1) * np.cos(y / 9.0) + 0.3 * np.sin(z / 4.0) + rng.normal(0, 0.01, shape)
    return f.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(19, 33, 47), (16, 16, 16), (1, 1, 300), (1, 40, 70), (35, 18, 20)])
def test_error_bound(oracle, dtype, shape):
    data = _smooth_field(shape, dtype, 1)
    for flags in ALL_FLAGS:
        for rel in (1e-2, 1e-4):
            rng_ = float(data.max()) - float(data.min())
            eb = rel * rng_
            qc, anchors = oracle.spline_construct(data, eb, rel, flags)
            rec = oracle.spline_reconstruct(qc, anchors, shape, eb, rel, flags)
            err = np.abs(rec.astype(np.float64) - data.astype(np.float64)).max()
            ulp = float(np.spacing(np.abs(data).max().astype(dtype)))
            assert err <= eb + 2 * ulp, (flags, rel, err / eb)


def test_alpha_schedule(oracle):
    """alpha is piecewise linear in rel_eb through (1e-1, 2), (1e-2, 1.75),
    (1e-3, 1.5), (1e-4, 1.25), (1e-5, 1) and flat outside
    (InterpolationPredictor.cu:2248-2271)."""
    pts = [(1e-1, 2.0), (1e-2, 1.75), (1e-3, 1.5), (1e-4, 1.25), (1e-5, 1.0)]
    for r, a in pts:
        assert oracle.alpha(r) == pytest.approx(a, abs=1e-15)
    for (r1, a1), (r0, a0) in zip(pts, pts[1:]):
        assert oracle.alpha(0.5 * (r0 + r1)) == pytest.approx(0.5 * (a0 + a1), rel=1e-12)
    assert oracle.alpha(0.7) == 2.0 and oracle.alpha(1e-9) == 1.0


def _point_level(shape):
    """Level of every grid point: k for unit 2^k, 4 for the anchors."""
    z, y, x = np.meshgrid(*[np.arange(d) for d in shape], indexing="ij")
    m = x | y | z | 16
    return np.log2(m & -m).astype(int)


@pytest.mark.parametrize("rel", [1e-1, 1e-3, 1e-6])
def test_level_bounds_follow_alpha_beta(oracle, rel):
    """Level k quantises with bound eb / min(alpha^k, beta = 4): on a noise
    field every error stays within it and the largest error of each level
    exceeds half of it, so a schedule that is off by a factor 2 (no beta
    clamp, alpha inverted, one level shifted) fails."""
    shape = (65, 65, 65)
    data = np.random.default_rng(5).uniform(-1, 1, shape)
    eb = rel * float(data.max() - data.min())
    flags = np.zeros(8, np.uint8)
    qc, an = oracle.spline_construct(data, eb, rel, flags)
    rec = oracle.spline_reconstruct(qc, an, shape, eb, rel, flags)
    err = np.abs(rec - data)
    lv = _point_level(shape)
    a = oracle.alpha(rel)
    assert not err[lv == 4].any()
    for k in range(4):
        bound = eb / min(a ** k, 4.0)
        e = err[lv == k]
        assert e.max() <= bound * (1 + 1e-12), (k, e.max() / bound)
        assert e.max() > 0.5 * bound, (k, e.max() / bound)


def test_delta_reconstruction(oracle):
    """Zero delta codes leave the old field unchanged; delta + base codes
    reconstruct close to the full codes (differences are rounding only)."""
    shape = (20, 33, 40)
    data = _smooth_field(shape, np.float64, 2)
    flags = np.array([1, 0, 1, 0, 0, 1, 0, 1], np.uint8)
    eb = 1e-5
    qc, anchors = oracle.spline_construct(data, eb, 1e-5, flags)
    full = oracle.spline_reconstruct(qc, anchors, shape, eb, 1e-5, flags)
    zero = np.zeros_like(qc)
    same = oracle.spline_delta_reconstruct(zero, full, eb, 1e-5, flags)
    assert np.array_equal(same, full)
    hi = qc.copy()
    hi[oracle.layout(40, 33, 20).n4:] //= 4
    base = oracle.spline_reconstruct(hi, anchors, shape, eb, 1e-5, flags)
    sumd = oracle.spline_delta_reconstruct(qc - hi, base, eb, 1e-5, flags)
    assert np.abs(sumd - full).max() < 1e-12


def _axis_field(axis, shape=(40, 40, 40), seed=3):
    rng = np.random.default_rng(seed)
    r = rng.normal(0, 1, max(shape)).astype(np.float64)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return r[{"x": x, "y": y, "z": z}[axis]]


def test_tuner_prefers_x_first_for_x_only_fields(oracle):
    """f = r[x]: only points predicted along x have an error, and it depends on
    x alone.  The x-first order predicts along x only the points with y and z
    on the coarser grid, a strict subset of the z-y-x order's set, and the
    multi-dimensional rule adds halved/thirded x errors on faces and cubes.
    So every level picks reverse (x first) and not the multi-dimensional rule."""
    scores, flags, rc = oracle.autotune(_axis_field("x"))
    assert rc == 0
    assert scores[1] < scores[0] and scores[4] < scores[3] and scores[13] < scores[12]
    assert list(flags) == [0, 0, 0, 0, 1, 1, 1, 1]


def test_tuner_prefers_z_first_for_z_only_fields(oracle):
    scores, flags, rc = oracle.autotune(_axis_field("z"))
    assert rc == 0
    assert scores[0] < scores[1] and scores[3] < scores[4] and scores[12] < scores[13]
    assert scores[0] < scores[2] and scores[12] < scores[14]
    assert list(flags) == [0] * 8


def test_tuner_undefined_for_small_axes(oracle):
    for shape in [(1, 1, 1000), (1, 50, 50), (16, 64, 64)]:
        scores, flags, rc = oracle.autotune(np.ones(shape, np.float32))
        assert rc == 1 and not flags.any() and not scores.any()


def test_pick_flags_tie_rules(oracle):
    assert list(oracle.pick_flags(np.zeros(15))) == [0] * 8
    s = np.arange(15, 0, -1, dtype=np.float64)  # later candidates always smaller
    f = oracle.pick_flags(s)
    # level 3: e1<e0 -> reverse, e2<e1 -> md; level 1: best 11 -> md, reverse; level 0: best 14 -> md
    assert list(f) == [1, 1, 1, 1, 1, 1, 1, 1]


# ---- closed-form pins of the tuner's sample geometry and of the rule
# selection in partial tiles (derivations in tests/closed_forms.py) ----

from closed_forms import (RULE_COEFS, RULE_SHAPES, TUNER_ORIGINS, TUNER_SHAPES, codes_by_point,  # noqa: E402
                          cubic_axes_field, expected_codes, noise_outside_windows, tuner_blocks)


def test_tuner_origin_table_is_the_reference_formula_by_hand():
    """The hand-derived origin table: spacing 128, the middle origin at
    d/2 (or d/2 - 128 k), the last window start at most d - 9."""
    for d, orig in TUNER_ORIGINS.items():
        assert all(b - a == 128 for a, b in zip(orig, orig[1:]))
        assert d // 2 in orig and orig[0] >= 8 and orig[-1] + 8 <= d - 1
        assert orig[0] - 128 < 8 and orig[-1] + 128 + 8 > d - 1
    assert len(tuner_blocks((384, 384, 256))) == 9
    assert len(tuner_blocks((512, 512, 512))) == 27


@pytest.mark.parametrize("shape_xyz", TUNER_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_tuner_reads_exactly_the_sample_windows(oracle, shape_xyz):
    """Noise outside the hand-derived windows and zeros inside: all 15 scores
    are exactly 0.  Then one impulse at the near and at the far corner of
    each window, on a zero field: every score is positive.  The far corner
    is local coordinate 16, or 8 where the window is cut by the field's end
    (a multiple of 8 is a grid point of every level, so every candidate
    sees it).  A window off by one point, an extra block or a missing block
    fails one of the two."""
    nx, ny, nz = shape_xyz
    dtype = np.float32
    scores, _, rc = oracle.autotune(noise_outside_windows(shape_xyz, dtype))
    assert rc == 0 and not scores.any(), scores
    f = np.zeros((nz, ny, nx), dtype)
    for ox, oy, oz in tuner_blocks(shape_xyz):
        far = [o + min(16, 8 * ((d - 1 - o) // 8)) for o, d in ((ox, nx), (oy, ny), (oz, nz))]
        for px, py, pz in ((ox, oy, oz), tuple(far)):
            f[pz, py, px] = 1.0
            scores, _, rc = oracle.autotune(f)
            f[pz, py, px] = 0.0
            assert rc == 0 and (scores > 0).all(), ((px, py, pz), scores)
    # just outside a window: one point before the first origin of x is read by no block
    ox, oy, oz = tuner_blocks(shape_xyz)[0]
    f[oz, oy, ox - 1] = 1.0
    scores, _, _ = oracle.autotune(f)
    assert not scores.any()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", RULE_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_partial_tile_rule_selection(oracle, dtype, shape):
    """Codes of f = cx x^3 + cy y^3 + cz z^3 at eb = 0.25 equal, point by
    point, 2 * the interpolation remainder of the rule that the closed-form
    availability test selects (cubic 0, quadratics -/+3u^3, linear -3gu^2,
    copy 3g^2u - 3gu^2 + u^3), for the four flag settings and a per-level
    mix; the reconstruction is exact."""
    nz, ny, nx = shape
    L = oracle.layout(nx, ny, nz)
    for coef in RULE_COEFS:
        data = cubic_axes_field(shape, coef, dtype)
        for flags in ALL_FLAGS + [np.array([1, 0, 1, 0, 0, 1, 0, 1], np.uint8)]:
            qc, an = oracle.spline_construct(data, 0.25, 0.0, flags)
            got = codes_by_point(L, qc, shape)
            want = expected_codes(shape, coef, flags)
            bad = np.argwhere(got != want)
            assert bad.size == 0, (coef, list(flags), bad[:4].tolist(),
                                   [(int(got[tuple(b)]), int(want[tuple(b)])) for b in bad[:4]])
            assert (want != 0).any()
            rec = oracle.spline_reconstruct(qc, an, shape, 0.25, 0.0, flags)
            assert np.array_equal(rec, data)


def test_cubic_polynomial_field_has_zero_codes_where_the_cubic_rule_applies(oracle):
    """A full trivariate polynomial of degree <= 3 in each variable (with
    cross terms) is predicted exactly by every cubic stencil and by the
    averages of cubic stencils, so every point whose used directions are all
    of order 4 has code 0; only points near a tile edge or the field end
    carry codes.  eb = 1/64 makes every error here (a multiple of 1/32)
    quantise exactly, so the reconstruction equals the field."""
    shape = (37, 35, 41)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(*[np.arange(d, dtype=np.float64) for d in shape], indexing="ij")
    f = x ** 3 - 2 * x * x * y + 3 * x * y * z - y ** 3 + 2 * z ** 3 - 5 * x * z * z + 7 * y + 11
    L = oracle.layout(nx, ny, nz)
    for flags in ALL_FLAGS:
        qc, an = oracle.spline_construct(f, 1.0 / 64, 0.0, flags)
        got = codes_by_point(L, qc, shape)
        # points all of whose neighbours at distance 3u exist inside their tile
        # on every axis: codes zero for every rule set
        from closed_forms import point_level
        lv = point_level(shape)
        interior = np.zeros(shape, bool)
        for k in range(4):
            u = 1 << k
            ok = lv == k
            for g, n in ((x, nx), (y, ny), (z, nz)):
                c = g.astype(int) % 16
                odd = (g.astype(int) % (2 * u)) == u
                ok &= ~odd | ((c >= 3 * u) & (c + 3 * u <= 16) & (g + 3 * u < n))
            interior |= ok
        assert not got[interior].any()
        assert got[~interior & (lv < 4)].any()
        rec = oracle.spline_reconstruct(qc, an, shape, 1.0 / 64, 0.0, flags)
        assert np.array_equal(rec, f)
