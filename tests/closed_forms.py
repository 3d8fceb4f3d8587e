"""Closed forms shared by the oracle pins and the GPU parity tests.

Two parts of the predictor had no pin beyond a restatement of the reference:
which 16^3 blocks the auto-tuner scores, and which interpolation rule a point
in the last, partial tile of an axis uses.  Both are fixed here by values
derived by hand from the reference's formulas, written out as numbers or as
polynomial error terms, never by calling the code under test.

Tuner sample geometry (InterpolationPredictor.cu:2277-2289), per axis of
length d, all integer arithmetic:
    mid = d / 2;  k = (mid - 8) / 128;  t = (d - 8 - 1 - mid) / 128
    origins = mid - 128 k, mid - 128 (k - 1), ..., mid + 128 t   (k + t + 1 of them)
and every block reads the 17^3 window [o, o + 16] of each axis, clipped to
the field.  Worked out by hand:
    d=17:   mid 8,   k 0, t 0            -> 8
    d=20:   mid 10,  k 0, t 1/128 = 0     -> 10
    d=40:   mid 20,  k 0, t 11/128 = 0    -> 20
    d=145:  mid 72,  k 0, t 64/128 = 0    -> 72
    d=256:  mid 128, k 0, t 119/128 = 0   -> 128
    d=272:  mid 136, k 1, t 127/128 = 0   -> 8, 136
    d=273:  mid 136, k 1, t 128/128 = 1   -> 8, 136, 264
    d=300:  mid 150, k 1, t 141/128 = 1   -> 22, 150, 278
    d=384:  mid 192, k 1, t 183/128 = 1   -> 64, 192, 320
    d=512:  mid 256, k 1, t 247/128 = 1   -> 128, 256, 384
    d=1000: mid 500, k 3, t 491/128 = 3   -> 116, 244, 372, 500, 628, 756, 884
so 384x384x256 has 3*3*1 = 9 blocks and 512^3 has 27.

Rule selection along one axis (1-D rule InterpolationPredictor.cu:1300-1326,
get_interp_order 1538-1545).  A point at tile-local coordinate c (global g)
of an axis of length N, predicted at unit u, has its neighbours at c - 3u,
c - u, c + u, c + 3u.  A neighbour is available when it lies in the tile's
17-point window [0, 16] and inside the field; c - u always is, and c + u
always lies in the window.  The rule is the highest-order one whose
neighbours are all available:
    order 4  cubic   (-1, 9, 9, -1)/16     c - 3u and c + 3u available
    order 3  left quadratic  (-1, 6, 3)/8  c - 3u and c + u, not c + 3u
    order 2  right quadratic (3, 6, -1)/8  c + 3u, not c - 3u
    order 1  linear  (1, 1)/2              c + u only
    order 0  copy of the c - u value       c + u outside the field
The reference states this as case1..case5 (case1 = "not the last tile",
inside which every window point is inside the field).

On the field f = a^3 along that axis (a = the axis coordinate), the error
f(g) - pred of each rule is, by the interpolation remainder of a cubic with
leading coefficient 1 (the product of (g - node) over the rule's nodes):
    cubic 0;  left quadratic (3u)(u)(-u) = -3u^3;  right quadratic
    (u)(-u)(-3u) = +3u^3;  linear f(g) - (f(g-u) + f(g+u))/2 = -3 g u^2;
    copy f(g) - f(g-u) = 3 g^2 u - 3 g u^2 + u^3.
A field constant along an axis is predicted exactly by every rule along it.

Stages (spline_layout_interpolate 1854-2025; a coordinate is "odd" at unit u
when it is an odd multiple of u): z-y-x order predicts z-odd points with x, y
even along z, then y-odd points with x even along y, then x-odd points along
x; the reverse order predicts x-odd points with y, z even along x, then
y-odd points with z even along y, then z-odd points along z.  The
multi-dimensional rule predicts points with one odd coordinate along it,
points with two odd coordinates by the face rule over those two directions
(first direction z, or y when z is even; second y or x), points with three
by the cube rule.  The face rule (1612-1667) ranks each direction's order by
class (cubic 3 > quadratic 2 > linear 1 > none 0), uses the direction of the
higher class, and averages the two when their classes tie; with both at
class 0 it predicts f(p - u e1) + f(p - u e2) - f(p - u e1 - u e2), exact on
fields that are sums of one-axis terms.  The cube rule (1695-1721) averages
the cubic predictions of the directions of order 4; with none, it uses x's
rule (left or right quadratic, linear, copy).

With eb = 0.25 (ebx2 = 0.5, eb_r = 4, alpha = 1) every error here is a
multiple of 0.5, so quantisation is exact: the code of a point is 2 * error
and every reconstructed value equals the original, so later predictions see
original values only.
"""
import numpy as np

TUNER_ORIGINS = {17: [8], 20: [10], 40: [20], 145: [72], 256: [128], 272: [8, 136], 273: [8, 136, 264],
                 300: [22, 150, 278], 384: [64, 192, 320], 512: [128, 256, 384],
                 1000: [116, 244, 372, 500, 628, 756, 884]}

# (nx, ny, nz) -> number of sample blocks, from the table above
TUNER_SHAPES = [(384, 384, 256), (300, 17, 273), (1000, 40, 272), (145, 20, 17), (512, 512, 512)]


def tuner_blocks(shape_xyz):
    nx, ny, nz = shape_xyz
    return [(ox, oy, oz) for oz in TUNER_ORIGINS[nz] for oy in TUNER_ORIGINS[ny] for ox in TUNER_ORIGINS[nx]]


def tuner_window_mask(shape_xyz):
    """[nz][ny][nx] bool: the points some sample block reads."""
    nx, ny, nz = shape_xyz
    m = np.zeros((nz, ny, nx), bool)
    for ox, oy, oz in tuner_blocks(shape_xyz):
        m[oz:oz + 17, oy:oy + 17, ox:ox + 17] = True
    return m


def noise_outside_windows(shape_xyz, dtype, seed=11):
    """Noise everywhere except on the sample windows, which hold 0: every
    score is exactly 0 if and only if the tuner reads no point outside the
    windows (a window shifted by one, or an extra block, reads noise)."""
    nx, ny, nz = shape_xyz
    rng = np.random.default_rng(seed)
    f = rng.uniform(1.0, 2.0, (nz, ny, nx)).astype(dtype)
    f[tuner_window_mask(shape_xyz)] = 0
    return f


def point_level(shape):
    """Level of every grid point of [nz][ny][nx]: k for unit 2^k, 4 for anchors."""
    z, y, x = np.meshgrid(*[np.arange(d) for d in shape], indexing="ij")
    m = x | y | z | 16
    return np.log2(m & -m).astype(int)


def codes_by_point(L, qc, shape):
    """Level-major codes (anchors first, then levels 3..0 from L.prefix, each
    in raster order) placed back on the [nz][ny][nx] grid; anchors read 0."""
    lv = point_level(shape).ravel()
    out = np.zeros(lv.size, np.int64)
    for k in range(4):
        pos = np.nonzero(lv == k)[0]
        out[pos] = qc[L.prefix[k]:L.prefix[k] + pos.size]
    return out.reshape(shape)


def _order(c, g, n, u):
    left3 = c >= 3 * u
    right1 = g + u < n
    right3 = (c + 3 * u <= 16) & (g + 3 * u < n)
    return np.where(left3 & right3, 4, np.where(left3 & right1, 3, np.where(right3, 2, np.where(right1, 1, 0))))


def _signature(order, g, u):
    g = g.astype(np.float64)
    return np.select([order == 4, order == 3, order == 2, order == 1],
                     [0.0, -3.0 * u ** 3, 3.0 * u ** 3, -3.0 * g * u * u],
                     3.0 * g * g * u - 3.0 * g * u * u + u ** 3)


def _cls(order):
    return np.select([order == 4, order >= 2, order == 1], [3, 2, 1], 0)


def cubic_axes_field(shape, coef, dtype):
    """f = cx x^3 + cy y^3 + cz z^3 on [nz][ny][nx] (coef = (cx, cy, cz))."""
    z, y, x = np.meshgrid(*[np.arange(d, dtype=np.float64) for d in shape], indexing="ij")
    return (coef[0] * x ** 3 + coef[1] * y ** 3 + coef[2] * z ** 3).astype(dtype)


def expected_codes(shape, coef, flags):
    """Codes of cubic_axes_field(shape, coef) at eb = 0.25, alpha = 1, from
    the closed forms of the module docstring."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    G = {"x": x, "y": y, "z": z}
    N = {"x": nx, "y": ny, "z": nz}
    lv = point_level(shape)
    err = np.zeros(shape, np.float64)
    for k in range(4):
        u = 1 << k
        md, rev = int(flags[k]), int(flags[4 + k])
        odd = {a: (G[a] % (2 * u)) == u for a in "xyz"}
        order = {a: _order(G[a] % 16, G[a], N[a], u) for a in "xyz"}
        sig = {a: c * _signature(order[a], G[a], u) for a, c in zip("xyz", coef)}
        ox, oy, oz = odd["x"], odd["y"], odd["z"]
        if md:
            nodd = ox.astype(int) + oy + oz
            e_line = np.where(ox, sig["x"], np.where(oy, sig["y"], sig["z"]))
            # face: first direction z (or y when z is even), second y or x
            o1 = np.where(~oz, order["y"], order["z"])
            s1 = np.where(~oz, sig["y"], sig["z"])
            o2 = np.where(~ox, order["y"], order["x"])
            s2 = np.where(~ox, sig["y"], sig["x"])
            c1, c2 = _cls(o1), _cls(o2)
            e_face = np.where((c1 == 0) & (c2 == 0), 0.0,
                              np.where(c1 == c2, (s1 + s2) / 2, np.where(c1 > c2, s1, s2)))
            anycubic = (order["x"] == 4) | (order["y"] == 4) | (order["z"] == 4)
            e_cube = np.where(anycubic, 0.0, sig["x"])
            e = np.select([nodd == 1, nodd == 2], [e_line, e_face], e_cube)
        elif rev:
            e = np.where(ox & ~oy & ~oz, sig["x"], np.where(oy & ~oz, sig["y"], sig["z"]))
        else:
            e = np.where(oz & ~ox & ~oy, sig["z"], np.where(oy & ~ox, sig["y"], sig["x"]))
        sel = lv == k
        err[sel] = e[sel]
    codes = 2 * err
    assert np.array_equal(codes, np.round(codes))
    return codes.astype(np.int64)


# (nz, ny, nx): every axis 16k + r with r in {1, 3, 5, 9, 15}, so each axis
# has interior tiles and a partial last tile of a different depth
RULE_SHAPES = [(21, 35, 41), (17, 47, 25), (31, 19, 35), (41, 25, 47)]
RULE_COEFS = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, -2, 1)]
