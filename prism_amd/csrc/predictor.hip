// predictor.hip -- multi-level cubic-spline prediction on gfx950.
//
// One workgroup (256 lanes = 4 wave64) per 16^3 tile.  The tile and its
// +1 halo (17^3 values of T) live in LDS; each of the 4 levels (unit 8,4,2,1)
// runs three stages, and every stage is a compile-time shape so the
// point <-> lane mapping uses constant divisors only.  Points of one stage
// read only points of earlier stages, so a stage is one parallel sweep
// followed by one barrier.  Arithmetic is in T (FP64 for double fields),
// evaluated in the reference's order with no FMA contraction, so the
// compressor and the decompressor reproduce identical values.
//
// Reference: InterpolationPredictor.cu -- stage maps 1786-1809 / 23-106,
// 1-D rule 1229-1352, multi-dimensional rule 1414-1750, level loop
// 1823-2026, tile staging 997-1170, compaction 2028-2088, auto-tuner 110-967.
#include "batch.hpp"
#include "common.hpp"
#include "extrema.hpp"

namespace pa {
namespace {

constexpr int kThreads = 256;
constexpr int kTileVals = kTS * kTS * kTS;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, 32 CUs each

enum Kind { BLUE, YELLOW, HOLLOW, HOLLOW_R, YELLOW_R, BLUE_R, LINE, FACE, CUBE };
enum Mode { COMPRESS, DECOMPRESS, TUNE };

__device__ __forceinline__ int sidx(int x, int y, int z) { return (z * kTS + y) * kTS + x; }

// Per-tile context.  IN = true for a tile that is not the last one along
// any axis: all 17^3 points lie inside the field and every boundary test of
// the stencils is decided at compile time.
template <typename T, bool IN = false>
struct Ctx {
    T* s;             // LDS tile
    int ox, oy, oz;   // tile origin
    bool c1x, c1y, c1z;  // tile is not the last one on the axis
    Geom g;
    T eb_r, ebx2;     // per-level values of the current level
    int32_t* qc_out;
    const int32_t* qc_in;
    bool q16;         // qc_in holds int16 codes (narrow hand-off of a progressive round)
    bool natural;     // tuner candidates 9..11
    double acc;       // tuner partial sum of this lane
    // the tile is not the last one along axis A (0 = x, 1 = y, 2 = z)
    template <int A>
    __device__ __forceinline__ bool c1() const {
        return IN || (A == 0 ? c1x : (A == 1 ? c1y : c1z));
    }
    // global coordinate v along axis A lies inside the field
    template <int A>
    __device__ __forceinline__ bool inf(int v) const {
        return IN || v < (A == 0 ? g.nx : (A == 1 ? g.ny : g.nz));
    }
};

// code i of the hand-off buffer (int32, or int16 when q16)
__device__ __forceinline__ int32_t code_at(const int32_t* q, bool q16, uint32_t i) {
    return q16 ? (int32_t)reinterpret_cast<const int16_t*>(q)[i] : q[i];
}

template <typename T>
__device__ __forceinline__ T cubic(T a, T b, T c, T d, bool natural) {
    if (natural) return ((((T)-3 * a + (T)23 * b) + (T)23 * c) - (T)3 * d) / (T)40;
    return (((-a + (T)9 * b) + (T)9 * c) - d) / (T)16;
}
template <typename T>
__device__ __forceinline__ T quad_l(T t0, T t1, T t2) {
    return ((-t0 + (T)6 * t1) + (T)3 * t2) / (T)8;
}
template <typename T>
__device__ __forceinline__ T quad_r(T t1, T t2, T t3) {
    return (((T)3 * t1 + (T)6 * t2) - t3) / (T)8;
}
template <typename T>
__device__ __forceinline__ T lin(T t1, T t2) {
    return (t1 + t2) / (T)2;
}

// get_interp_order (InterpolationPredictor.cu:1538-1545)
template <int U>
__device__ __forceinline__ int interp_order(int c, int g, int N, bool c1) {
    const int b = (c >= 3 * U) ? 3 : 1;
    const int f = ((c + 3 * U <= 16) && (c1 || (g + 3 * U < N))) ? 3 : ((c1 || (g + U < N)) ? 1 : 0);
    return (b == 3) ? ((f == 3) ? 4 : ((f == 1) ? 3 : 0)) : ((f == 3) ? 2 : ((f == 1) ? 1 : 0));
}

// 1-D rule along axis A (0 = x, 1 = y, 2 = z), InterpolationPredictor.cu:1269-1326
template <typename T, int U, int A, bool IN>
__device__ __forceinline__ T pred_1d(const Ctx<T, IN>& t, int x, int y, int z) {
    const int c = A == 0 ? x : (A == 1 ? y : z);
    const int o = A == 0 ? t.ox : (A == 1 ? t.oy : t.oz);
    const int N = A == 0 ? t.g.nx : (A == 1 ? t.g.ny : t.g.nz);
    const bool c1 = t.template c1<A>();
    const int gc = o + c;
    const int i0 = c - 3 * U >= 0 ? c - 3 * U : 0;
    const int i1 = c - U >= 0 ? c - U : 0;
    const int i2 = c + U < kTS ? c + U : 0;
    const int i3 = c + 3 * U < kTS ? c + 3 * U : 0;
    const int st = A == 0 ? 1 : (A == 1 ? kTS : kTS * kTS);
    const int base = sidx(x, y, z) - c * st;
    const T t0 = t.s[base + i0 * st], t1 = t.s[base + i1 * st];
    const T t2 = t.s[base + i2 * st], t3 = t.s[base + i3 * st];
    const bool c2 = c >= 3 * U, c3 = c + 3 * U <= 16, c4 = gc + 3 * U < N, c5 = gc + U < N;
    T pred = t1;
    if ((c1 && !c2 && !c3) || (!c1 && !c2 && !(c3 && c4) && c5))
        pred = lin(t1, t2);
    else if ((c1 && !c2 && c3) || (!c1 && !c2 && c3 && c4))
        pred = quad_r(t1, t2, t3);
    else if ((c1 && c2 && !c3) || (!c1 && c2 && !(c3 && c4) && c5))
        pred = quad_l(t0, t1, t2);
    else if ((c1 && c2 && c3) || (!c1 && c2 && c3 && c4))
        pred = cubic(t0, t1, t2, t3, t.natural);
    return pred;
}

template <int U>
__device__ __forceinline__ void taps(int c, int (&id)[4]) {
    id[0] = (c - 3 * U >= 0) ? c - 3 * U : 0;
    id[1] = (c - U >= 0) ? c - U : 0;
    id[2] = (c + U <= 16) ? c + U : 0;
    id[3] = (c + 3 * U <= 16) ? c + 3 * U : 0;
}

// FACE rule, InterpolationPredictor.cu:1546-1669
template <typename T, int U, bool IN>
__device__ __forceinline__ T pred_face(const Ctx<T, IN>& t, int x, int y, int z) {
    int idz[4], idy[4], idx[4];
    taps<U>(z, idz);
    taps<U>(y, idy);
    taps<U>(x, idx);
    T s1[4], s2[4], pred;
    int i1, i2;
    if ((x % (2 * U)) == 0) {  // directions z (1) and y (2)
#pragma unroll
        for (int i = 0; i < 4; ++i) s1[i] = t.s[sidx(x, y, idz[i])];
#pragma unroll
        for (int i = 0; i < 4; ++i) s2[i] = t.s[sidx(x, idy[i], z)];
        pred = t.s[sidx(x, idy[1], idz[1])];
        i1 = interp_order<U>(z, t.oz + z, t.g.nz, t.template c1<2>());
        i2 = interp_order<U>(y, t.oy + y, t.g.ny, t.template c1<1>());
    } else if ((y % (2 * U)) == 0) {  // directions z (1) and x (2)
#pragma unroll
        for (int i = 0; i < 4; ++i) s1[i] = t.s[sidx(x, y, idz[i])];
#pragma unroll
        for (int i = 0; i < 4; ++i) s2[i] = t.s[sidx(idx[i], y, z)];
        pred = t.s[sidx(idx[1], y, idz[1])];
        i1 = interp_order<U>(z, t.oz + z, t.g.nz, t.template c1<2>());
        i2 = interp_order<U>(x, t.ox + x, t.g.nx, t.template c1<0>());
    } else {  // directions y (1) and x (2)
#pragma unroll
        for (int i = 0; i < 4; ++i) s1[i] = t.s[sidx(x, idy[i], z)];
#pragma unroll
        for (int i = 0; i < 4; ++i) s2[i] = t.s[sidx(idx[i], y, z)];
        pred = t.s[sidx(idx[1], idy[1], z)];
        i1 = interp_order<U>(y, t.oy + y, t.g.ny, t.template c1<1>());
        i2 = interp_order<U>(x, t.ox + x, t.g.nx, t.template c1<0>());
    }
    const bool nat = t.natural;
    if (i1 == 4 && i2 == 4) {
        pred = (cubic(s1[0], s1[1], s1[2], s1[3], nat) + cubic(s2[0], s2[1], s2[2], s2[3], nat)) / (T)2;
    } else if (i1 != 4 && i2 == 4) {
        pred = cubic(s2[0], s2[1], s2[2], s2[3], nat);
    } else if (i1 == 4 && i2 != 4) {
        pred = cubic(s1[0], s1[1], s1[2], s1[3], nat);
    } else if (i1 == 3 && i2 == 3) {
        pred = quad_l(s2[0], s2[1], s2[2]);
        pred += quad_l(s1[0], s1[1], s1[2]);
        pred /= (T)2;
    } else if (i1 == 3 && i2 == 2) {
        pred = quad_r(s2[1], s2[2], s2[3]);
        pred += quad_l(s1[0], s1[1], s1[2]);
        pred /= (T)2;
    } else if (i1 == 3 && i2 < 2) {
        pred = quad_l(s1[0], s1[1], s1[2]);
    } else if (i1 == 2 && i2 == 3) {
        pred = quad_r(s1[1], s1[2], s1[3]);
        pred += quad_l(s2[0], s2[1], s2[2]);
        pred /= (T)2;
    } else if (i1 == 2 && i2 == 2) {
        pred = quad_r(s1[1], s1[2], s1[3]);
        pred += quad_r(s2[1], s2[2], s2[3]);
        pred /= (T)2;
    } else if (i1 == 2 && i2 < 2) {
        pred = quad_r(s1[1], s1[2], s1[3]);
    } else if (i1 <= 1 && i2 == 3) {
        pred = quad_l(s2[0], s2[1], s2[2]);
    } else if (i1 <= 1 && i2 == 2) {
        pred = quad_r(s2[1], s2[2], s2[3]);
    } else if (i1 == 1 && i2 == 1) {
        pred = lin(s2[1], s2[2]);
        pred += lin(s1[1], s1[2]);
        pred /= (T)2;
    } else if (i1 == 1 && i2 < 1) {
        pred = lin(s1[1], s1[2]);
    } else if (i1 == 0 && i2 == 1) {
        pred = lin(s2[1], s2[2]);
    } else {
        pred = (s1[1] + s2[1]) - pred;
    }
    return pred;
}

// CUBE rule, InterpolationPredictor.cu:1671-1723
template <typename T, int U, bool IN>
__device__ __forceinline__ T pred_cube(const Ctx<T, IN>& t, int x, int y, int z) {
    int idz[4], idy[4], idx[4];
    taps<U>(z, idz);
    taps<U>(y, idy);
    taps<U>(x, idx);
    const int iz = interp_order<U>(z, t.oz + z, t.g.nz, t.template c1<2>());
    const int iy = interp_order<U>(y, t.oy + y, t.g.ny, t.template c1<1>());
    const int ix = interp_order<U>(x, t.ox + x, t.g.nx, t.template c1<0>());
    T tx[4], ty[4], tz[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) tx[i] = t.s[sidx(idx[i], y, z)];
#pragma unroll
    for (int i = 0; i < 4; ++i) tz[i] = iz == 4 ? t.s[sidx(x, y, idz[i])] : (T)0;
#pragma unroll
    for (int i = 0; i < 4; ++i) ty[i] = iy == 4 ? t.s[sidx(x, idy[i], z)] : (T)0;
    const bool nat = t.natural;
    const T px1 = cubic(tx[0], tx[1], tx[2], tx[3], nat);
    const T py1 = cubic(ty[0], ty[1], ty[2], ty[3], nat);
    const T pz1 = cubic(tz[0], tz[1], tz[2], tz[3], nat);
    T pred = tx[1];
    if (iz == 4 && iy == 4 && ix == 4) pred = ((px1 + py1) + pz1) / (T)3;
    else if (iz == 4 && iy == 4) pred = (pz1 + py1) / (T)2;
    else if (iz == 4 && ix == 4) pred = (pz1 + px1) / (T)2;
    else if (iy == 4 && ix == 4) pred = (py1 + px1) / (T)2;
    else if (iz == 4) pred = pz1;
    else if (iy == 4) pred = py1;
    else if (ix == 4) pred = px1;
    else if (ix == 3) pred = quad_l(tx[0], tx[1], tx[2]);
    else if (ix == 2) pred = quad_r(tx[1], tx[2], tx[3]);
    else if (ix == 1) pred = lin(tx[1], tx[2]);
    return pred;
}

// Stage geometry: point p (0 <= p < total) of stage K at unit U, in the
// reference's thread-visiting order (it defines the tuner's lane assignment
// p mod 256; the other modes do not depend on it).
template <int U, int K>
__host__ __device__ constexpr int stage_total() {
    constexpr int N = 16 / (2 * U);
    if constexpr (K <= BLUE_R) {
        constexpr int XC = (K == HOLLOW || K == HOLLOW_R) ? N : ((K == YELLOW_R || K == BLUE_R) ? 2 * N + 1 : N + 1);
        constexpr int YC = (K == YELLOW || K == YELLOW_R) ? N : ((K == HOLLOW || K == BLUE_R) ? 2 * N + 1 : N + 1);
        constexpr int ZC = (K == BLUE || K == BLUE_R) ? N : ((K == YELLOW || K == HOLLOW) ? 2 * N + 1 : N + 1);
        return XC * YC * ZC;
    } else if constexpr (K == LINE) {
        return 3 * N * (N + 1) * (N + 1);
    } else if constexpr (K == FACE) {
        return 3 * N * N * (N + 1);
    } else {
        return N * N * N;
    }
}

template <int U, int K>
__host__ __device__ constexpr int stage_iters() {
    return (stage_total<U, K>() + kThreads - 1) / kThreads;
}

template <int U, int K>
__device__ __forceinline__ void stage_point(int p, int& x, int& y, int& z) {
    constexpr int N = 16 / (2 * U);
    if constexpr (K <= BLUE_R) {
        constexpr int X0 = (K == HOLLOW || K == HOLLOW_R) ? U : 0;
        constexpr int XS = (K == YELLOW_R || K == BLUE_R) ? U : 2 * U;
        constexpr int XC = (K == HOLLOW || K == HOLLOW_R) ? N : ((K == YELLOW_R || K == BLUE_R) ? 2 * N + 1 : N + 1);
        constexpr int Y0 = (K == YELLOW || K == YELLOW_R) ? U : 0;
        constexpr int YS = (K == HOLLOW || K == BLUE_R) ? U : 2 * U;
        constexpr int YC = (K == YELLOW || K == YELLOW_R) ? N : ((K == HOLLOW || K == BLUE_R) ? 2 * N + 1 : N + 1);
        constexpr int Z0 = (K == BLUE || K == BLUE_R) ? U : 0;
        constexpr int ZS = (K == YELLOW || K == HOLLOW) ? U : 2 * U;
        x = X0 + XS * (p % XC);
        y = Y0 + YS * ((p / XC) % YC);
        z = Z0 + ZS * (p / (XC * YC));
    } else if constexpr (K == LINE) {
        constexpr int L = N * (N + 1) * (N + 1), Q = (N + 1) * (N + 1);
        const int grp = p / L, m = p % L;
        const int i = m / Q, j = (m % Q) / (N + 1), k = (m % Q) % (N + 1);
        if (grp == 0) { x = 2 * U * i + U; y = 2 * U * j; z = 2 * U * k; }
        else if (grp == 1) { x = 2 * U * k; y = 2 * U * i + U; z = 2 * U * j; }
        else { x = 2 * U * j; y = 2 * U * k; z = 2 * U * i + U; }
    } else if constexpr (K == FACE) {
        constexpr int L = N * N * (N + 1), Q = N * N;
        const int grp = p / L, m = p % L;
        const int i = m / Q, j = (m % Q) / N, k = (m % Q) % N;
        if (grp == 0) { x = 2 * U * i; y = 2 * U * j + U; z = 2 * U * k + U; }
        else if (grp == 1) { x = 2 * U * k + U; y = 2 * U * i; z = 2 * U * j + U; }
        else { x = 2 * U * j + U; y = 2 * U * k + U; z = 2 * U * i; }
    } else {
        constexpr int Q = N * N;
        const int i = p / Q, j = (p % Q) / N, k = (p % Q) % N;
        x = 2 * U * i + U;
        y = 2 * U * j + U;
        z = 2 * U * k + U;
    }
}

// Whether point (x, y, z) of the tile is predicted in this sweep: inside the
// tile (the +1 halo included for inclusive stages; the last stage of the
// tile row keeps its far plane) and inside the field.
template <typename T, bool IN>
__device__ __forceinline__ bool visited(const Ctx<T, IN>& t, int x, int y, int z, bool inclusive) {
    const bool in = inclusive ? (x <= 16 && y <= 16 && z <= 16)
                              : (x < 16 && y < 16 && z < 16 + (t.template c1<2>() ? 0 : 1));
    return in && t.template inf<0>(t.ox + x) && t.template inf<1>(t.oy + y) && t.template inf<2>(t.oz + z);
}

template <typename T, int U, int K, bool IN>
__device__ __forceinline__ T predict(const Ctx<T, IN>& t, int x, int y, int z) {
    if constexpr (K == BLUE || K == BLUE_R) {
        return pred_1d<T, U, 2>(t, x, y, z);
    } else if constexpr (K == YELLOW || K == YELLOW_R) {
        return pred_1d<T, U, 1>(t, x, y, z);
    } else if constexpr (K == HOLLOW || K == HOLLOW_R) {
        return pred_1d<T, U, 0>(t, x, y, z);
    } else if constexpr (K == LINE) {
        if ((z % (2 * U)) > 0) return pred_1d<T, U, 2>(t, x, y, z);
        if ((y % (2 * U)) > 0) return pred_1d<T, U, 1>(t, x, y, z);
        return pred_1d<T, U, 0>(t, x, y, z);
    } else if constexpr (K == FACE) {
        return pred_face<T, U>(t, x, y, z);
    } else {
        return pred_cube<T, U>(t, x, y, z);
    }
}

template <int U>
constexpr int level_of() {
    return U == 1 ? 0 : (U == 2 ? 1 : (U == 4 ? 2 : 3));
}

// The lane id through an opaque move: index arithmetic derived from it in
// one phase of the kernel is not merged with that of another phase and kept
// live in registers between them.
__device__ __forceinline__ int lane_id() {
    int v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "v"((int)threadIdx.x));
    __builtin_assume(v >= 0 && v < kThreads);
    return v;
}

// Barrier for LDS producer/consumer hand-off between stages: waits for this
// wave's LDS operations only, so outstanding global loads (prefetched codes)
// and stores (emitted codes) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Quantise v against pred, write the reconstruction back and return the
// code (InterpolationPredictor.cu:1327-1341).
template <typename T>
__device__ __forceinline__ int32_t quantize_val(T& v, T pred, T eb_r, T ebx2) {
    const T err = v - pred;
    T code = (err < (T)0 ? -err : err) * eb_r + (T)1;
    code = err < (T)0 ? -code : code;
    const int qv = (int)(code / (T)2);
    code = (T)qv;
    v = pred + code * ebx2;
    return qv;
}

template <typename T, bool IN>
__device__ __forceinline__ int32_t quantize(const Ctx<T, IN>& t, T& v, T pred) {
    return quantize_val(v, pred, t.eb_r, t.ebx2);
}

// Decompression first fills the LDS slot of every point with the bit
// pattern of its own code (anchors hold their value, points of skipped
// levels 0); a stage then reads the slot of the point it reconstructs and
// overwrites it with pred + code * ebx2.  Predictions read only points of
// earlier stages, which already hold values.
__device__ __forceinline__ float code_slot(int32_t q, float) { return __int_as_float(q); }
__device__ __forceinline__ double code_slot(int32_t q, double) {
    return __longlong_as_double((long long)(uint32_t)q);
}
__device__ __forceinline__ int32_t slot_code(float v) { return __float_as_int(v); }
__device__ __forceinline__ int32_t slot_code(double v) { return (int32_t)(uint32_t)__double_as_longlong(v); }

// Where a compression stage puts its codes: straight to the level-major
// array (levels 1-3), into registers with their LDS slots (the first two
// stages of level 0, whose values later stages still read), or into the
// point's own LDS slot (the last stage of level 0: no later stage reads its
// values).  Level 0's codes then leave the tile as whole rows.
enum Sink { TO_GLOBAL, TO_REGS, TO_SLOT };
constexpr int kRegCodes = 8;  // points per lane of a level-0 stage (line or point form)
struct RegCodes {
    int32_t code[kRegCodes];
    int32_t slot[kRegCodes];  // -1: no own point
};

// ----------------------------------------------------------- point form --
// Lane l visits points p = l + 256 * it of the stage (tuner, and the
// multi-dimensional stages).
template <typename T, int U, int K, int M, bool IN, int SINK = TO_GLOBAL>
__device__ __forceinline__ void stage(Ctx<T, IN>& t, bool inclusive, RegCodes* rc = nullptr) {
    constexpr int TOTAL = stage_total<U, K>();
    constexpr int LEVEL = level_of<U>();
    static_assert(SINK != TO_REGS || stage_iters<U, K>() <= kRegCodes, "register codes: at most 8 points per lane");
    const int lane = lane_id();
    if constexpr (SINK == TO_REGS) {
#pragma unroll
        for (int k = 0; k < kRegCodes; ++k) rc->slot[k] = -1;
    }
#pragma unroll
    for (int it = 0; it < stage_iters<U, K>(); ++it) {
        const int p = lane + it * kThreads;
        if (TOTAL % kThreads != 0 && p >= TOTAL) break;
        int x, y, z;
        stage_point<U, K>(p, x, y, z);
        if (!visited(t, x, y, z, inclusive)) continue;
        const T pred = predict<T, U, K>(t, x, y, z);
        T& v = t.s[sidx(x, y, z)];
        if constexpr (M == TUNE) {
            const T e = v - pred;
            t.acc += (double)(e < (T)0 ? -e : e);
        } else if constexpr (M == COMPRESS) {
            const int32_t qv = quantize(t, v, pred);
            const bool own = x < 16 && y < 16 && z < 16;
            if constexpr (SINK == TO_GLOBAL) {
                if (own) t.qc_out[lm_index<LEVEL>(t.g, t.ox + x, t.oy + y, t.oz + z)] = qv;
            } else if constexpr (SINK == TO_REGS) {
                rc->code[it] = qv;
                rc->slot[it] = own ? sidx(x, y, z) : -1;
            } else {
                v = code_slot(qv, T());
            }
        } else {
            v = pred + (T)slot_code(v) * t.ebx2;
        }
    }
    lds_barrier();
}

// ------------------------------------------------------------ line form --
// 1-D stages along y or z run line by line: lane l owns one line along the
// prediction axis A (consecutive lanes are consecutive lines along x, so
// their codes are neighbours in memory), reads the N + 1 known values of the
// line once and predicts its N points with compile-time stencils.  Which
// stencil a point uses depends only on its position along A and on the
// tile's place in the field, so the choice is the same for every lane.
// The values produced are those of the per-point rule above (same formula,
// same operation order); only the lane assignment differs, and points of one
// stage never read each other.
enum Stencil { S_KEEP, S_LIN, S_QUAD_R, S_QUAD_L, S_CUBIC };

template <int U, int K, bool EXCL>
struct LineShape {
    static constexpr int N = 16 / (2 * U);
    static constexpr int A = (K == BLUE || K == BLUE_R) ? 2 : ((K == YELLOW || K == YELLOW_R) ? 1 : 0);
    // P: first other axis (x, or y when A = x); Q: second (z, or y when A = z)
    static constexpr bool FULL_P = (K == HOLLOW || K == YELLOW_R || K == BLUE_R);
    static constexpr bool FULL_Q = (K == YELLOW || K == HOLLOW || K == BLUE_R);
    static constexpr int SP = FULL_P ? U : 2 * U;
    static constexpr int SQ = FULL_Q ? U : 2 * U;
    // the exclusive last stage of level 0 visits x < 16, y < 16, z < 16 (its
    // z = 16 plane would lie outside the field: only the last tile along z
    // may visit it, and there oz + 16 >= nz)
    static constexpr int CP = EXCL ? 16 : (FULL_P ? 2 * N + 1 : N + 1);
    static constexpr int CQ = EXCL ? 16 : (FULL_Q ? 2 * N + 1 : N + 1);
    static constexpr int LINES = CP * CQ;
    static constexpr int ITERS = (LINES + kThreads - 1) / kThreads;
    static constexpr int SLOTS = ITERS * N;
    static constexpr int STA = A == 0 ? 1 : (A == 1 ? kTS : kTS * kTS);
};

// Stencil of point j of a line (position c = (2j+1)U along the axis), the
// rule of InterpolationPredictor.cu:1269-1326 with c2 = (j >= 1) and
// c3 = (j <= N - 2).
template <int U>
__device__ __forceinline__ int stencil_of(int j, int N, bool c1, int o, int nax) {
    const int gc = o + (2 * j + 1) * U;
    const bool c2 = j >= 1, c3 = j <= N - 2, c4 = gc + 3 * U < nax, c5 = gc + U < nax;
    if ((c1 && !c2 && !c3) || (!c1 && !c2 && !(c3 && c4) && c5)) return S_LIN;
    if ((c1 && !c2 && c3) || (!c1 && !c2 && c3 && c4)) return S_QUAD_R;
    if ((c1 && c2 && !c3) || (!c1 && c2 && !(c3 && c4) && c5)) return S_QUAD_L;
    if ((c1 && c2 && c3) || (!c1 && c2 && c3 && c4)) return S_CUBIC;
    return S_KEEP;
}

template <typename T>
__device__ __forceinline__ T apply_stencil(int k, T t0, T t1, T t2, T t3) {
    switch (k) {
        case S_LIN: return lin(t1, t2);
        case S_QUAD_R: return quad_r(t1, t2, t3);
        case S_QUAD_L: return quad_l(t0, t1, t2);
        case S_CUBIC: return cubic(t0, t1, t2, t3, false);
        default: return t1;
    }
}

struct LineLane {
    int base;       // LDS index of the line's c = 0 point
    bool on;        // line inside the tile's sweep and the field
    bool own;       // line inside the 16^3 tile proper (codes are stored)
    uint32_t idx0;  // level-major index of point j = 0
    uint32_t st;    // level-major stride between consecutive points
};

template <typename T, int U, int K, bool EXCL, bool IN>
__device__ __forceinline__ LineLane line_lane(const Ctx<T, IN>& t, int it) {
    using S = LineShape<U, K, EXCL>;
    LineLane r;
    const int l = lane_id() + it * kThreads;
    const int ip = l % S::CP, iq = l / S::CP;
    const int pc = S::SP * ip, qc = S::SQ * iq;
    int x, y, z;  // point j = 0
    if constexpr (S::A == 0) { x = U; y = pc; z = qc; }
    else if constexpr (S::A == 1) { x = pc; y = U; z = qc; }
    else { x = pc; y = qc; z = U; }
    const int gx = t.ox + x, gy = t.oy + y, gz = t.oz + z;
    bool on = (S::LINES % kThreads == 0 || l < S::LINES);
    if constexpr (S::A != 0) on = on && t.template inf<0>(gx);
    if constexpr (S::A != 1) on = on && t.template inf<1>(gy);
    if constexpr (S::A != 2) on = on && t.template inf<2>(gz);
    r.on = on;
    r.own = x < 16 && y < 16 && z < 16;
    r.base = sidx(x, y, z) - U * S::STA;
    constexpr int LEVEL = level_of<U>();
    r.idx0 = lm_index<LEVEL>(t.g, gx, gy, gz);
    r.st = 0;
    if constexpr (S::N >= 2) {
        // the index is affine along the line: the varying coordinate keeps
        // its parity and its halves grow by one per point
        const uint32_t i1 = S::A == 0 ? lm_index<LEVEL>(t.g, gx + 2 * U, gy, gz)
                                     : (S::A == 1 ? lm_index<LEVEL>(t.g, gx, gy + 2 * U, gz)
                                                  : lm_index<LEVEL>(t.g, gx, gy, gz + 2 * U));
        r.st = i1 - r.idx0;
    }
    return r;
}

template <typename T, int U, int K, bool EXCL, int M, bool IN, int SINK = TO_GLOBAL>
__device__ __forceinline__ void line_stage(Ctx<T, IN>& t, RegCodes* rc = nullptr) {
    using S = LineShape<U, K, EXCL>;
    constexpr int N = S::N;
    static_assert(SINK != TO_REGS || S::ITERS * N <= kRegCodes, "register codes: at most 8 points per lane");
    if constexpr (SINK == TO_REGS) {
#pragma unroll
        for (int k = 0; k < kRegCodes; ++k) rc->slot[k] = -1;
    }
    const int o = S::A == 0 ? t.ox : (S::A == 1 ? t.oy : t.oz);
    const int nax = S::A == 0 ? t.g.nx : (S::A == 1 ? t.g.ny : t.g.nz);
    const bool c1 = t.template c1<S::A>();
#pragma unroll
    for (int it = 0; it < S::ITERS; ++it) {
        const LineLane ln = line_lane<T, U, K, EXCL>(t, it);
        if (!ln.on) continue;
        T k[N + 1];
#pragma unroll
        for (int j = 0; j <= N; ++j) k[j] = t.s[ln.base + 2 * j * U * S::STA];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (!t.template inf<S::A>(o + (2 * j + 1) * U)) break;  // beyond the field; so are later points
            const T pred = apply_stencil(stencil_of<U>(j, N, c1, o, nax), k[j > 0 ? j - 1 : 0], k[j], k[j + 1],
                                         k[j + 2 <= N ? j + 2 : 0]);
            const int slot = ln.base + (2 * j + 1) * U * S::STA;
            T& v = t.s[slot];
            if constexpr (M == COMPRESS) {
                const int32_t qv = quantize(t, v, pred);
                if constexpr (SINK == TO_GLOBAL) {
                    if (ln.own) t.qc_out[ln.idx0 + j * ln.st] = qv;
                } else if constexpr (SINK == TO_REGS) {
                    rc->code[it * N + j] = qv;
                    rc->slot[it * N + j] = ln.own ? slot : -1;
                } else {
                    v = code_slot(qv, T());
                }
            } else {
                v = pred + (T)slot_code(v) * t.ebx2;
            }
        }
    }
    lds_barrier();
}

// ---------------------------------------------------------------- levels --
// The three stages of one level; outside tuning the last stage of level 0
// uses the exclusive border (InterpolationPredictor.cu:1940-2025).
template <typename T, bool IN>
__device__ __forceinline__ void level0_codes_out(Ctx<T, IN>& t, const RegCodes& a, const RegCodes& b);

template <typename T, int U, int M, bool IN>
__device__ __forceinline__ void level(Ctx<T, IN>& t, bool md, bool rev) {
    constexpr bool EX = M != TUNE && U == 1;
    if constexpr (M == COMPRESS && U == 1) {
        RegCodes a, b;
        if (md) {
            stage<T, 1, LINE, M, IN, TO_REGS>(t, true, &a);
            stage<T, 1, FACE, M, IN, TO_REGS>(t, true, &b);
            stage<T, 1, CUBE, M, IN, TO_SLOT>(t, false);
        } else if (rev) {
            line_stage<T, 1, HOLLOW_R, false, M, IN, TO_REGS>(t, &a);
            line_stage<T, 1, YELLOW_R, false, M, IN, TO_REGS>(t, &b);
            line_stage<T, 1, BLUE_R, true, M, IN, TO_SLOT>(t);
        } else {
            line_stage<T, 1, BLUE, false, M, IN, TO_REGS>(t, &a);
            line_stage<T, 1, YELLOW, false, M, IN, TO_REGS>(t, &b);
            line_stage<T, 1, HOLLOW, true, M, IN, TO_SLOT>(t);
        }
        level0_codes_out(t, a, b);
        return;
    }
    if (md) {
        stage<T, U, LINE, M>(t, true);
        stage<T, U, FACE, M>(t, true);
        stage<T, U, CUBE, M>(t, !EX);
    } else if constexpr (M == TUNE) {
        if (rev) {
            stage<T, U, HOLLOW_R, M>(t, true);
            stage<T, U, YELLOW_R, M>(t, true);
            stage<T, U, BLUE_R, M>(t, true);
        } else {
            stage<T, U, BLUE, M>(t, true);
            stage<T, U, YELLOW, M>(t, true);
            stage<T, U, HOLLOW, M>(t, true);
        }
    } else if (rev) {
        line_stage<T, U, HOLLOW_R, false, M>(t);
        line_stage<T, U, YELLOW_R, false, M>(t);
        line_stage<T, U, BLUE_R, EX, M>(t);
    } else {
        line_stage<T, U, BLUE, false, M>(t);
        line_stage<T, U, YELLOW, false, M>(t);
        line_stage<T, U, HOLLOW, EX, M>(t);
    }
}

// Decompression start: the slot of every tile point (17^3, halo included)
// gets its code, its anchor or 0.  Level L's points are the multiples of
// U = 2^L in the tile that are not all multiples of 2U.  Along a row (y, z)
// of level L the level-major index is affine in x (the compaction formula,
// InterpolationPredictor.cu:2070-2083): with i = x / U,
//   rows with y/U or z/U odd hold every i:  idx = base_L(y, z) + ox/U + i,
//   other rows hold the odd i only:          idx = base_L(y, z) + ox/(2U) + i/2.
// The 404 row bases of a tile (17^2 + 9^2 + 5^2 + 3^2) are computed once
// into LDS; then every code load is one table read and an add, and all
// loads are issued before the first slot is written.
constexpr int kRows0 = 289, kRows1 = 81, kRows2 = 25, kRows3 = 9;
constexpr int kRowTab = kRows0 + kRows1 + kRows2 + kRows3;  // 404
template <int L>
constexpr int row_off() {
    return L == 0 ? 0 : (L == 1 ? kRows0 : (L == 2 ? kRows0 + kRows1 : kRows0 + kRows1 + kRows2));
}

template <int L>
__device__ __forceinline__ int32_t row_base(const Geom& g, int gy, int gz) {
    // 32-bit unsigned arithmetic, exact below code_len < 2^31 (as lm_index)
    const uint32_t b = (uint32_t)gy >> L, c = (uint32_t)gz >> L;
    uint32_t r = b * (uint32_t)g.lx[L] + c * (uint32_t)g.lxy[L] + (uint32_t)g.prefix[L] -
                 ((c + 1) >> 1) * (uint32_t)g.lxy[L + 1];
    if ((c & 1) == 0) r -= ((b + 1) >> 1) * (uint32_t)g.lx[L + 1];
    return (int32_t)r;
}

// Level 0's codes of the 16^3 tile, all in their LDS slots once the
// register-held ones join them, written out row by row: lane (x, y) =
// (l mod 16, l / 16) takes z = 0..15; a row (y, z) with y or z odd holds
// every x (16 adjacent codes), the others the odd x (8 adjacent codes).
// Row bases advance by 2 lxy0 - lxy1 from z to z + 2.
template <typename T, bool IN>
__device__ __forceinline__ void level0_codes_out(Ctx<T, IN>& t, const RegCodes& a, const RegCodes& b) {
#pragma unroll
    for (int k = 0; k < kRegCodes; ++k) {
        if (a.slot[k] >= 0) t.s[a.slot[k]] = code_slot(a.code[k], T());
        if (b.slot[k] >= 0) t.s[b.slot[k]] = code_slot(b.code[k], T());
    }
    lds_barrier();
    const int lane = lane_id();
    const int x = lane & 15, y = lane >> 4;
    const int gx = t.ox + x, gy = t.oy + y;
    if (!(t.template inf<0>(gx) && t.template inf<1>(gy))) return;
    const uint32_t step2 = (uint32_t)(2 * t.g.lxy[0] - t.g.lxy[1]);
    const uint32_t base_e = (uint32_t)row_base<0>(t.g, gy, t.oz), base_o = (uint32_t)row_base<0>(t.g, gy, t.oz + 1);
    const uint32_t off_full = (uint32_t)gx, off_half = (uint32_t)((t.ox >> 1) + (x >> 1));
#pragma unroll
    for (int z = 0; z < 16; ++z) {
        if (!t.template inf<2>(t.oz + z)) break;
        const bool full = ((y | z) & 1) != 0;
        if (full || (x & 1)) {
            const uint32_t base = ((z & 1) ? base_o : base_e) + (uint32_t)(z >> 1) * step2;
            t.qc_out[base + (full ? off_full : off_half)] = slot_code(t.s[sidx(x, y, z)]);
        }
    }
}

template <typename T, int L, bool IN>
__device__ __forceinline__ void row_entry(const Ctx<T, IN>& t, int32_t* tab, int row) {
    constexpr int U = 1 << L, E = 16 / U + 1;
    const int gy = t.oy + U * (row % E), gz = t.oz + U * (row / E);
    tab[row_off<L>() + row] = (t.template inf<1>(gy) && t.template inf<2>(gz)) ? row_base<L>(t.g, gy, gz) : 0;
}

template <typename T, bool IN>
__device__ __forceinline__ void fill_rows(const Ctx<T, IN>& t, int32_t* tab) {
    const int lane = lane_id();
    row_entry<T, 0>(t, tab, lane);
    if (lane + kThreads < kRows0) row_entry<T, 0>(t, tab, lane + kThreads);
    if (lane < kRows1) row_entry<T, 1>(t, tab, lane);
    else if (lane < kRows1 + kRows2) row_entry<T, 2>(t, tab, lane - kRows1);
    else if (lane < kRows1 + kRows2 + kRows3) row_entry<T, 3>(t, tab, lane - kRows1 - kRows2);
}

// Level 0: lane l keeps x = l mod 16 and takes rows r = l / 16 + 16 it; the
// x = 16 column is a second, short pass.  Levels 1-3: lane l takes lattice
// point p = l + 256 it of the level's E^3 lattice (x fastest).
constexpr int kFill0Iters = (kRows0 + 15) / 16;                   // 19
constexpr int kFill0EdgeIters = (kRows0 + kThreads - 1) / kThreads;  // 2
template <int L>
struct FillShape {
    static constexpr int U = 1 << L, E = 16 / U + 1, TOTAL = E * E * E;
    static constexpr int ITERS = (TOTAL + kThreads - 1) / kThreads;
};

// Row r0 + 16 it of the 17 x 17 rows of a tile (r0 = lane / 16 < 16):
// (y, z) = (r0 - it, it), or (r0 - it + 17, it - 1) when r0 < it (16 is
// -1 mod 17), and (16, 16) for the last row (it = 18, r0 = 0 only).
struct RowYZ {
    int y, z;
    bool valid;
};
template <int IT>
__device__ __forceinline__ RowYZ row_yz(int r0) {
    RowYZ w;
    if constexpr (IT == kFill0Iters - 1) {
        w.y = 16;
        w.z = 16;
        w.valid = r0 == 0;
    } else {
        const bool wrap = r0 < IT;
        w.y = r0 - IT + (wrap ? kTS : 0);
        w.z = IT - (wrap ? 1 : 0);
        w.valid = true;
    }
    return w;
}

template <typename T, int IT, bool IN>
__device__ __forceinline__ void fill0_one(const Ctx<T, IN>& t, const int32_t* tab, bool load, int r0, int x, bool xin,
                                          uint32_t off_full, uint32_t off_half, int32_t& v) {
    const RowYZ w = row_yz<IT>(r0);
    const bool full = ((w.y | w.z) & 1) != 0;
    v = 0;
    if (load && w.valid && xin && (full || (x & 1)) && t.template inf<1>(t.oy + w.y) && t.template inf<2>(t.oz + w.z))
        v = code_at(t.qc_in, t.q16, (uint32_t)tab[r0 + 16 * IT] + (full ? off_full : off_half));
}

template <typename T, int IT, bool IN>
__device__ __forceinline__ void fill0_iter(const Ctx<T, IN>& t, const int32_t* tab, bool load, int r0, int x,
                                           bool xin, uint32_t off_full, uint32_t off_half,
                                           int32_t (&v)[kFill0Iters]) {
    if constexpr (IT < kFill0Iters) {
        fill0_one<T, IT>(t, tab, load, r0, x, xin, off_full, off_half, v[IT]);
        fill0_iter<T, IT + 1>(t, tab, load, r0, x, xin, off_full, off_half, v);
    }
}

template <typename T, bool IN>
__device__ __forceinline__ void fill0_load(const Ctx<T, IN>& t, const int32_t* tab, bool load,
                                           int32_t (&v)[kFill0Iters], int32_t (&e)[kFill0EdgeIters]) {
    const int lane = lane_id();
    const int x = lane & 15, r0 = lane >> 4;
    const bool xin = t.template inf<0>(t.ox + x);
    fill0_iter<T, 0>(t, tab, load, r0, x, xin, (uint32_t)(t.ox + x), (uint32_t)((t.ox >> 1) + (x >> 1)), v);
#pragma unroll
    for (int it = 0; it < kFill0EdgeIters; ++it) {  // x = 16: level 0 where y or z is odd
        const int r = lane + kThreads * it;
        const int y = r % kTS, z = r / kTS;
        e[it] = 0;
        if (load && r < kRows0 && ((y | z) & 1) && t.template inf<0>(t.ox + 16) && t.template inf<1>(t.oy + y) &&
            t.template inf<2>(t.oz + z))
            e[it] = code_at(t.qc_in, t.q16, (uint32_t)tab[r] + (uint32_t)(t.ox + 16));
    }
}

template <typename T, int IT>
__device__ __forceinline__ void fill0_put(T* s, int r0, int x, const int32_t (&v)[kFill0Iters]) {
    if constexpr (IT < kFill0Iters) {
        const RowYZ w = row_yz<IT>(r0);
        if (w.valid && (((w.y | w.z) & 1) || (x & 1))) s[(r0 + 16 * IT) * kTS + x] = code_slot(v[IT], T());
        fill0_put<T, IT + 1>(s, r0, x, v);
    }
}

template <typename T>
__device__ __forceinline__ void fill0_store(T* s, const int32_t (&v)[kFill0Iters],
                                            const int32_t (&e)[kFill0EdgeIters]) {
    const int lane = lane_id();
    const int x = lane & 15, r0 = lane >> 4;
    fill0_put<T, 0>(s, r0, x, v);
#pragma unroll
    for (int it = 0; it < kFill0EdgeIters; ++it) {
        const int r = lane + kThreads * it;
        const int y = r % kTS, z = r / kTS;
        if (r < kRows0 && ((y | z) & 1)) s[r * kTS + 16] = code_slot(e[it], T());
    }
}

template <typename T, int L, bool IN>
__device__ __forceinline__ void fill_load(const Ctx<T, IN>& t, const int32_t* tab, bool load,
                                          int32_t (&v)[FillShape<L>::ITERS]) {
    using F = FillShape<L>;
    const int lane = lane_id();
#pragma unroll
    for (int it = 0; it < F::ITERS; ++it) {
        const int p = lane + it * kThreads;
        const int i = p % F::E, row = p / F::E, j = row % F::E, k = row / F::E;
        const bool full = ((j | k) & 1) != 0;
        const bool mine = (F::TOTAL % kThreads == 0 || p < F::TOTAL) && (full || (i & 1));
        v[it] = 0;
        if (load && mine && t.template inf<0>(t.ox + i * F::U) && t.template inf<1>(t.oy + j * F::U) &&
            t.template inf<2>(t.oz + k * F::U))
            v[it] = code_at(t.qc_in, t.q16,
                            (uint32_t)(tab[row_off<L>() + row] + (full ? (t.ox >> L) + i : (t.ox >> (L + 1)) + (i >> 1))));
    }
}

template <typename T, int L>
__device__ __forceinline__ void fill_store(T* s, const int32_t (&v)[FillShape<L>::ITERS]) {
    using F = FillShape<L>;
    const int lane = lane_id();
#pragma unroll
    for (int it = 0; it < F::ITERS; ++it) {
        const int p = lane + it * kThreads;
        const int i = p % F::E, row = p / F::E, j = row % F::E, k = row / F::E;
        if ((F::TOTAL % kThreads == 0 || p < F::TOTAL) && ((i | j | k) & 1))
            s[sidx(i * F::U, j * F::U, k * F::U)] = code_slot(v[it], T());
    }
}

template <typename T>
struct LevelEb {
    T eb_r[4], ebx2[4];
};

// Level bounds from a device block laid out as {absolute eb, range,
// eb_r[4] and ebx2[4] in float64, then eb_r[4] and ebx2[4] in float32}.
template <typename T>
__device__ __forceinline__ LevelEb<T> load_level_ebs(const double* e) {
    LevelEb<T> le;
    if constexpr (sizeof(T) == 4) {
        const float* q = reinterpret_cast<const float*>(e + 10);
        for (int i = 0; i < 4; ++i) {
            le.eb_r[i] = q[i];
            le.ebx2[i] = q[4 + i];
        }
    } else {
        for (int i = 0; i < 4; ++i) {
            le.eb_r[i] = e[2 + i];
            le.ebx2[i] = e[6 + i];
        }
    }
    return le;
}

// Tile plus halo into LDS: all global loads of a lane are issued before the
// first LDS write so the tile costs one memory latency.  Lane l takes
// x = l mod 16 of the rows r = l / 16 + 16 it (row r = 17 z + y); since
// 16 = 17 - 1, row r0 + 16 it (it < 18) is (y, z) = (r0 - it, it) when
// r0 >= it and (r0 - it + 17, it - 1) otherwise, so row positions and field
// offsets are a compile-time step plus one select per row; the last row
// (288, taken by r0 = 0 at it = 18) is (16, 16).  The x = 16 column is a
// second, short pass.  Field offsets below 2^31 (the layout's limit) are
// 32-bit.
template <typename T, bool IN>
__device__ __forceinline__ void load_tile(const Ctx<T, IN>& t, const T* data) {
    const int lane = lane_id();
    const int x = lane & 15, r0 = lane >> 4;
    const uint32_t nx = (uint32_t)t.g.nx, nxy = nx * (uint32_t)t.g.ny;
    const bool xin = t.template inf<0>(t.ox + x);
    const T* base = data + ((uint32_t)t.oz * nxy + (uint32_t)t.oy * nx + (uint32_t)t.ox);
    const uint32_t lane_off = (uint32_t)r0 * nx + (uint32_t)x;
    const uint32_t step = nxy - nx, wrap_off = 17 * nx - nxy;  // per row step; back by 17 rows of y
    T v[kFill0Iters], e[kFill0EdgeIters];
#pragma unroll
    for (int it = 0; it < kFill0Iters; ++it) {
        const bool last = it == kFill0Iters - 1, wrap = r0 < it;
        const int y = last ? 16 : r0 - it + (wrap ? kTS : 0), z = last ? 16 : it - (wrap ? 1 : 0);
        const uint32_t off =
            last ? 16 * nx + 16 * nxy + (uint32_t)x : lane_off + (uint32_t)it * step + (wrap ? wrap_off : 0u);
        v[it] = (T)0;
        if ((it < kFill0Iters - 1 || r0 == 0) && xin && t.template inf<1>(t.oy + y) && t.template inf<2>(t.oz + z))
            v[it] = base[off];
    }
#pragma unroll
    for (int it = 0; it < kFill0EdgeIters; ++it) {  // x = 16
        const int r = lane + kThreads * it;
        const int y = r % kTS, z = r / kTS;
        e[it] = (T)0;
        if (r < kRows0 && t.template inf<0>(t.ox + 16) && t.template inf<1>(t.oy + y) && t.template inf<2>(t.oz + z))
            e[it] = base[(uint32_t)y * nx + (uint32_t)z * nxy + 16u];
    }
#pragma unroll
    for (int it = 0; it < kFill0Iters; ++it)
        if (it < kFill0Iters - 1 || r0 == 0) t.s[(r0 + 16 * it) * kTS + x] = v[it];
#pragma unroll
    for (int it = 0; it < kFill0EdgeIters; ++it) {
        const int r = lane + kThreads * it;
        if (r < kRows0) t.s[r * kTS + 16] = e[it];
    }
}

// Zero the code-array entries the predictor does not write: the anchor
// slots, the pads that 8-align each level region and the level-0 tail.
struct Gaps {
    int64_t lo[5], hi[5];
};

__device__ __forceinline__ void zero_gaps(int32_t* p, const int64_t* lo, const int64_t* hi, int64_t blk,
                                          int64_t nblk) {
    for (int r = 0; r < 5; ++r)
        for (int64_t i = lo[r] + blk * blockDim.x + threadIdx.x; i < hi[r]; i += nblk * blockDim.x) p[i] = 0;
}

// The decoded bit-planes of a progressive round (device layout: level l,
// plane k at base[l] + k * stride[l]) and the round's window, for a
// reconstruction that reads its codes from them.
struct PlaneSrc {
    const uint8_t* bp;
    const int32_t* begin;  // per level lr = 3 - l, planes [begin, end)
    const int32_t* end;
    int64_t base[4], stride[4];
    int64_t size;  // bytes of bp that may be read
};

template <typename T>
struct SplineArgs {
    const T* data;
    Geom g;
    LevelEb<T> le;
    const double* le_dev;  // when set, the level bounds come from this device block instead of `le`
    // when set, the device word picks where the codes come from: 0 int32
    // codes in qc_in, 1 int16 codes in qc_in, 2 the bit-planes of `ps`
    const int32_t* narrow;
    bool q16;
    const uint8_t* flags8;
    const int32_t* d_maxlevel;
    int32_t* qc_out;
    const int32_t* qc_in;
    T* anchors;
    const T* anchors_in;
    const T* old;
    T* out;
    Gaps gaps;    // compression: the gaps of qc_out, zeroed by the same launch
    int gaps_on;  // (its workgroups share them out; disjoint from the codes)
    PlaneSrc ps;  // decompression: the round's bit-planes (ps.bp null: none)
    bool planes;  // resolved: the codes come from ps
};

// Compression launches clear the code array's gaps alongside the tiles.
template <typename T, int M>
__device__ __forceinline__ void clear_gaps(const SplineArgs<T>& a) {
    if constexpr (M == COMPRESS) {
        if (a.gaps_on) zero_gaps(a.qc_out, a.gaps.lo, a.gaps.hi, blockIdx.x, gridDim.x);
    }
}

// Per-tile context of tile (bx, by, bz).
template <typename T, bool IN>
__device__ __forceinline__ Ctx<T, IN> tile_ctx(const SplineArgs<T>& a, T* tile, int bx, int by, int bz, bool q16) {
    const Geom& g = a.g;
    Ctx<T, IN> t;
    t.s = tile;
    t.ox = bx * 16;
    t.oy = by * 16;
    t.oz = bz * 16;
    t.c1x = t.ox + 16 < g.nx;
    t.c1y = t.oy + 16 < g.ny;
    t.c1z = t.oz + 16 < g.nz;
    t.g = g;
    t.qc_out = a.qc_out;
    t.qc_in = a.qc_in;
    t.q16 = q16;
    t.natural = false;
    t.acc = 0.0;
    return t;
}

// ------------------------------------------------------- decompression --
// A tile's codes and corner anchors in registers, between their loads and
// their LDS slots.
template <typename T>
struct TileCodes {
    int32_t v3[FillShape<3>::ITERS], v2[FillShape<2>::ITERS], v1[FillShape<1>::ITERS], v0[kFill0Iters],
        e0[kFill0EdgeIters];
    T anc;
    int aslot;
};

// Issues the loads of every code of the tile (levels up to maxlevel) from the
// row table of the tile in LDS, and of its eight corner anchors
// (x_reset_scratch_data 1051-1061).
template <typename T, bool IN>
__device__ __forceinline__ void codes_load(const Ctx<T, IN>& t, const int32_t* rowtab, int maxlevel,
                                           const T* anchors_in, int bx, int by, int bz, TileCodes<T>& c) {
    const Geom& g = t.g;
    fill_load<T, 3>(t, rowtab, 3 <= maxlevel, c.v3);
    fill_load<T, 2>(t, rowtab, 2 <= maxlevel, c.v2);
    fill_load<T, 1>(t, rowtab, 1 <= maxlevel, c.v1);
    fill0_load<T>(t, rowtab, 0 <= maxlevel, c.v0, c.e0);
    c.anc = (T)0;
    c.aslot = -1;
    if (threadIdx.x < 8) {
        const int cx = threadIdx.x & 1, cy = (threadIdx.x >> 1) & 1, cz = threadIdx.x >> 2;
        const int ax = bx + cx, ay = by + cy, az = bz + cz;
        c.aslot = sidx(16 * cx, 16 * cy, 16 * cz);
        if (anchors_in && ax < g.gx && ay < g.gy && az < g.gz)
            c.anc = anchors_in[ax + (int64_t)ay * g.gx + (int64_t)az * g.gx * g.gy];
    }
}

template <typename T>
__device__ __forceinline__ void codes_store(T* tile, const TileCodes<T>& c) {
    fill_store<T, 3>(tile, c.v3);
    fill_store<T, 2>(tile, c.v2);
    fill_store<T, 1>(tile, c.v1);
    fill0_store<T>(tile, c.v0, c.e0);
    if (c.aslot >= 0) tile[c.aslot] = c.anc;
}

// ------------------------------------------ codes from the bit-planes --
// A progressive round whose window is narrow reads each tile's codes from
// the decoded bit-planes instead of a code hand-off: the inverse shuffle
// (bitplane.cu:645-667) folded into the reconstruction's code load
// (InterpolationPredictor.cu:2209-2239).  Lane l takes rows l and 256 + l of
// the tile's 404 code rows, in the row table's order.  A row's codes are
// adjacent in its level's segment, so on each plane of the window they lie
// in at most 3 adjacent bytes (17 codes from any bit offset): one 8-byte
// load from the 4-byte aligned address below them, a funnel shift, and the
// bit order reversed within each byte put code m of the run at bit m.  Code
// m gathers bit m of each window plane k at bit 31 - k, and the negabinary
// inverse of that word is the round's code, as the inverse shuffle
// computes it.
struct PlaneRow {
    int row;     // row (j, k) = (row mod E, row / E) of the level's lattice
    bool full;   // the row holds every i, or the odd i only
    int nv;      // codes of the row inside the field, read
    int sh, np;  // bit offset of code 0 in its byte; planes read
    uint32_t w[kPlaneWindow];
};

// 8 bytes at a 4-byte aligned address in one load
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef u32x2 u32x2_a4 __attribute__((aligned(4)));

template <int L, typename T, bool IN>
__device__ __forceinline__ void plane_row_load(const Ctx<T, IN>& t, const PlaneSrc& ps, int row, int maxlevel, int b,
                                               int e, PlaneRow& r) {
    constexpr int U = 1 << L, E = 16 / U + 1;
    const int j = row % E, k = row / E;
    r.row = row;
    r.full = ((j | k) & 1) != 0;
    const int n = r.full ? E : (E - 1) / 2;
    r.nv = 0;
    if (L <= maxlevel && t.template inf<1>(t.oy + U * j) && t.template inf<2>(t.oz + U * k)) {
        if (IN) {
            r.nv = n;
        } else {  // i = m (full) or 2m + 1 (odd i) with ox + U i < nx
            const int ilim = (t.g.nx - t.ox + U - 1) >> L;
            r.nv = min(n, r.full ? ilim : ilim >> 1);
        }
    }
    r.np = r.nv > 0 && e > b ? e - b : 0;
    r.sh = 0;
#pragma unroll
    for (int q = 0; q < kPlaneWindow; ++q) r.w[q] = 0u;
    if (r.np == 0) return;
    // the run's first code, relative to the level's segment (32-bit: the
    // code array is shorter than 2^31); the row base computed here, so the
    // loads need no row table
    const uint32_t start = (uint32_t)row_base<L>(t.g, t.oy + U * j, t.oz + U * k) +
                           (uint32_t)(r.full ? t.ox >> L : t.ox >> (L + 1));
    const uint32_t rel = start - (uint32_t)t.g.prefix[L];
    r.sh = (int)(rel & 7u);
    const uint32_t byte = rel >> 3;
    const int64_t stride = ps.stride[L];
    const uint8_t* p0 = ps.bp + ps.base[L] + (byte & ~3u) + (int64_t)b * stride;
    // the low word lies inside plane k's row of bytes (stride is a multiple
    // of 8 and byte < stride); the high one may pass its end, never the buffer's
    const uint8_t* last = ps.bp + ps.size - 8;
#pragma unroll
    for (int q = 0; q < kPlaneWindow; ++q) {
        if (q < r.np) {
            const uint8_t* p = p0 + q * stride;
            uint32_t lo, hi = 0u;
            if (p <= last) {
                const u32x2 v = *reinterpret_cast<const u32x2_a4*>(p);
                lo = v.x;
                hi = v.y;
            } else {
                lo = *reinterpret_cast<const uint32_t*>(p);
            }
            r.w[q] = __builtin_amdgcn_alignbyte(hi, lo, byte & 3u);
        }
    }
}

template <int L, typename T>
__device__ __forceinline__ void plane_row_store(T* s, const PlaneRow& r, int b) {
    constexpr int U = 1 << L, E = 16 / U + 1;
    const int j = r.row % E, k = r.row / E;
    // the digits of codes 4c .. 4c + 3 in bytes 0..3 of dg[c], digit q (plane
    // b + q) at bit 7 - q of its byte: x * 0x204081 puts bit m of a 4-bit x
    // at bit 8m (its four shifted copies do not overlap, so nothing carries)
    constexpr int G = (E + 3) / 4;
    uint32_t dg[G];
#pragma unroll
    for (int c = 0; c < G; ++c) dg[c] = 0u;
#pragma unroll
    for (int q = 0; q < kPlaneWindow; ++q) {
        if (q < r.np) {
            const uint32_t w = __builtin_bitreverse32(__builtin_bswap32(r.w[q])) >> r.sh;
#pragma unroll
            for (int c = 0; c < G; ++c) dg[c] |= ((((w >> (4 * c)) & 0xFu) * 0x204081u) & 0x01010101u) << (7 - q);
        }
    }
    // digit q belongs at bit 31 - b - q: the code's byte to the top, then
    // right by b (b < 32 wherever a plane was read; else the byte is 0)
    const int sb = b & 31;
    const int n = r.full ? E : (E - 1) / 2;
    const int first = sidx(r.full ? 0 : U, U * j, U * k), step = r.full ? U : 2 * U;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        if (m < n) {
            const uint32_t top = __builtin_amdgcn_perm(0u, dg[m >> 2], 0x000C0C0Cu | ((uint32_t)(m & 3) << 24));
            s[first + step * m] = code_slot(m < r.nv ? nb_inv(top >> sb) : 0, T());
        }
    }
}

// Every code slot and corner anchor of the tile from the bit-planes, in two
// halves (the loads, then the slot writes) so that other loads can be
// issued between them.  The first rows of all lanes are level 0's; the
// second rows are the rest of level 0, then levels 1-3 (each level's rows
// with its lattice sizes known at compile time).
template <typename T>
struct PlaneCodes {
    PlaneRow r0, r1;
    int wb[4];  // per level lr = 3 - L
    T anc;
    int aslot;
};

template <typename T, bool IN>
__device__ __forceinline__ void plane_codes_load(const Ctx<T, IN>& t, const PlaneSrc& ps, int maxlevel,
                                                 const T* anchors_in, int bx, int by, int bz, PlaneCodes<T>& c) {
    const Geom& g = t.g;
    int we[4];
#pragma unroll
    for (int lr = 0; lr < 4; ++lr) {
        c.wb[lr] = ps.begin[lr];
        we[lr] = ps.end[lr];
    }
    const int lane = lane_id();
    constexpr int k1 = kRows0 + kRows1, k2 = k1 + kRows2;
    const int rg = lane + kThreads;
    plane_row_load<0>(t, ps, lane, maxlevel, c.wb[3], we[3], c.r0);
    if (rg < kRows0) plane_row_load<0>(t, ps, rg, maxlevel, c.wb[3], we[3], c.r1);
    else if (rg < k1) plane_row_load<1>(t, ps, rg - kRows0, maxlevel, c.wb[2], we[2], c.r1);
    else if (rg < k2) plane_row_load<2>(t, ps, rg - k1, maxlevel, c.wb[1], we[1], c.r1);
    else if (rg < kRowTab) plane_row_load<3>(t, ps, rg - k2, maxlevel, c.wb[0], we[0], c.r1);
    c.anc = (T)0;
    c.aslot = -1;
    if (threadIdx.x < 8) {
        const int cx = threadIdx.x & 1, cy = (threadIdx.x >> 1) & 1, cz = threadIdx.x >> 2;
        const int ax = bx + cx, ay = by + cy, az = bz + cz;
        c.aslot = sidx(16 * cx, 16 * cy, 16 * cz);
        if (anchors_in && ax < g.gx && ay < g.gy && az < g.gz)
            c.anc = anchors_in[ax + (int64_t)ay * g.gx + (int64_t)az * g.gx * g.gy];
    }
}

template <typename T, bool IN>
__device__ __forceinline__ void plane_codes_store(const Ctx<T, IN>& t, const PlaneCodes<T>& c) {
    constexpr int k1 = kRows0 + kRows1, k2 = k1 + kRows2;
    const int rg = lane_id() + kThreads;
    plane_row_store<0>(t.s, c.r0, c.wb[3]);
    if (rg < kRows0) plane_row_store<0>(t.s, c.r1, c.wb[3]);
    else if (rg < k1) plane_row_store<1>(t.s, c.r1, c.wb[2]);
    else if (rg < k2) plane_row_store<2>(t.s, c.r1, c.wb[1]);
    else if (rg < kRowTab) plane_row_store<3>(t.s, c.r1, c.wb[0]);
    if (c.aslot >= 0) t.s[c.aslot] = c.anc;
}

// The previous reconstruction of the tile (delta form); lane: (x, y) =
// (l mod 16, l / 16), iteration: z.
constexpr int kOutIters = 4096 / kThreads;
template <typename T, bool IN>
__device__ __forceinline__ void prev_load(const Ctx<T, IN>& t, const T* old, T (&prev)[kOutIters]) {
    const Geom& g = t.g;
    const int lane = lane_id();
    const int gx = t.ox + (lane & 15), gy = t.oy + (lane >> 4);
    const bool in_xy = t.template inf<0>(gx) && t.template inf<1>(gy);
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const T* src = old + (int64_t)gx + (int64_t)gy * g.nx + (int64_t)t.oz * nxy;
#pragma unroll
    for (int z = 0; z < kOutIters; ++z) {
        prev[z] = (T)0;
        // read once, here only: no cache allocation (reconstruction 158 -> 154 us)
        if (in_xy && t.template inf<2>(t.oz + z)) prev[z] = __builtin_nontemporal_load(&src[z * nxy]);
    }
}

// The reconstructed tile (plus the previous one in the delta form) to `out`.
template <typename T, bool IN>
__device__ __forceinline__ void tile_out(const Ctx<T, IN>& t, bool delta, const T (&prev)[kOutIters], T* out) {
    const Geom& g = t.g;
    const int lane = lane_id();
    const int x = lane & 15, y = lane >> 4;
    const int gx = t.ox + x, gy = t.oy + y;
    const bool in_xy = t.template inf<0>(gx) && t.template inf<1>(gy);
    const int64_t nxy = (int64_t)g.nx * g.ny;
    T* dst = out + (int64_t)gx + (int64_t)gy * g.nx + (int64_t)t.oz * nxy;
#pragma unroll
    for (int z = 0; z < kOutIters; ++z) {
        if (in_xy && t.template inf<2>(t.oz + z)) {
            const T v = t.s[sidx(x, y, z)];
            // written once and read only by a later round: no cache allocation
            __builtin_nontemporal_store(delta ? v + prev[z] : v, &dst[z * nxy]);
        }
    }
}

// The four levels of a tile, coarsest first.
template <typename T, int M, bool IN>
__device__ __forceinline__ void tile_levels(Ctx<T, IN>& t, const LevelEb<T>& le, const uint8_t (&fl)[8], int maxlevel) {
    if (3 <= maxlevel) { t.eb_r = le.eb_r[3]; t.ebx2 = le.ebx2[3]; level<T, 8, M>(t, fl[3], fl[7]); }
    if (2 <= maxlevel) { t.eb_r = le.eb_r[2]; t.ebx2 = le.ebx2[2]; level<T, 4, M>(t, fl[2], fl[6]); }
    if (1 <= maxlevel) { t.eb_r = le.eb_r[1]; t.ebx2 = le.ebx2[1]; level<T, 2, M>(t, fl[1], fl[5]); }
    if (0 <= maxlevel) { t.eb_r = le.eb_r[0]; t.ebx2 = le.ebx2[0]; level<T, 1, M>(t, fl[0], fl[4]); }
}

// One tile: compress (load, predict, quantise, store codes) or decompress
// (fill the code slots, reconstruct, add old, store).
// Where a decompressing tile finds its codes (resolved per launch, one
// kernel body each): int32 or int16 codes of a hand-off array, or the
// round's decoded bit-planes.
enum CodeSrc { CODES_32, CODES_16, CODES_PLANES };

// DELTA: a delta round (a.old given), a compile-time fact so that the
// previous values' loads are unconditional in interior tiles and the code
// slots wait for the code loads alone
template <typename T, int M, bool IN, int CS = CODES_32, bool DELTA = false>
__device__ __forceinline__ void tile_run(const SplineArgs<T>& a, T* tile, int32_t* rowtab, int bx, int by, int bz) {
    const Geom& g = a.g;
    Ctx<T, IN> t = tile_ctx<T, IN>(a, tile, bx, by, bz, CS == CODES_16);
    uint8_t fl[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) fl[i] = a.flags8[i];
    const int maxlevel = a.d_maxlevel ? *a.d_maxlevel : 4;

    // decompression: every code of the tile and the previous reconstruction
    // (delta form) are requested before any is needed, the codes first:
    // loads return in issue order, so the code slots are filled while the
    // previous values (needed only at the output) are still in flight
    T prev[kOutIters];
    if constexpr (M == COMPRESS) {
        load_tile(t, a.data);
        __syncthreads();
        if (threadIdx.x == 0) a.anchors[bx + (int64_t)by * g.gx + (int64_t)bz * g.gx * g.gy] = tile[0];
    } else {
        // a raised wave priority while the tile's loads are issued puts them in
        // flight ahead of other workgroups' stage arithmetic (reconstruction
        // 1-2 us faster)
        __builtin_amdgcn_s_setprio(3);
        if constexpr (CS != CODES_PLANES) {  // the plane loads compute their row bases themselves
            fill_rows(t, rowtab);
            lds_barrier();
        }
        if constexpr (CS == CODES_PLANES) {
            PlaneCodes<T> c;
            plane_codes_load(t, a.ps, maxlevel, a.anchors_in, bx, by, bz, c);
            if constexpr (DELTA) prev_load(t, a.old, prev);
            __builtin_amdgcn_s_setprio(0);
            plane_codes_store(t, c);
        } else {
            TileCodes<T> c;
            codes_load(t, rowtab, maxlevel, a.anchors_in, bx, by, bz, c);
            if constexpr (DELTA) prev_load(t, a.old, prev);
            __builtin_amdgcn_s_setprio(0);
            codes_store(tile, c);
        }
        __syncthreads();
    }

    tile_levels<T, M>(t, a.le, fl, maxlevel);

    if constexpr (M == DECOMPRESS) tile_out(t, DELTA, prev, a.out);
}

// a decompressing tile of either form: delta (a.old given) or anchored
template <typename T, bool IN, int CS>
__device__ __forceinline__ void tile_run_dec(const SplineArgs<T>& a, T* tile, int32_t* rowtab, int bx, int by,
                                             int bz) {
    if (a.old) tile_run<T, DECOMPRESS, IN, CS, true>(a, tile, rowtab, bx, by, bz);
    else tile_run<T, DECOMPRESS, IN, CS, false>(a, tile, rowtab, bx, by, bz);
}

// ------------------------------------------------------------ 2-D fields --
// A field with nz = 1 (and ny > 1) holds only the plane z = 0 of each tile.
// Of the stages only these have points inside it: YELLOW / YELLOW_R
// (y-lines), HOLLOW / HOLLOW_R (x-lines), and in the multi-dimensional form
// the LINE points along x and y and the FACE points with z = 0; every z
// stage is empty.  One wave owns one tile's 17 x 17 plane in LDS (z = 0 of
// the tile indexing, so the per-point rules above apply unchanged) and runs
// the stages of each level in the same order, with the same inclusive /
// exclusive sweeps, over 64 lanes; points of one stage never read each
// other, so only the order of stages matters and a wave-level fence
// separates them.  A workgroup holds four tiles.
constexpr int kPlaneWaves = kThreads / 64;  // tiles per workgroup
constexpr int kPlaneVals = kTS * kTS;       // 289

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

enum PlaneKind { P_YLINE, P_XLINE, P_FACE };

// Points of one stage of level U on the plane: y-lines at x = 0, S, 2S, ..
// <= 16 (points y = (2j+1)U), x-lines at y = 0, S, .. (points x = (2j+1)U;
// y < 16 for the exclusive last stage of level 0), or the face points.
template <typename T, int U, int KIND, int S, bool EXCL, int M>
__device__ __forceinline__ void plane_stage(Ctx<T>& t, int lane) {
    constexpr int N = 16 / (2 * U);
    constexpr int LINES = EXCL ? 16 : 16 / S + 1;
    constexpr int TOTAL = KIND == P_FACE ? N * N : LINES * N;
    constexpr int LEVEL = level_of<U>();
#pragma unroll
    for (int it = 0; it < (TOTAL + 63) / 64; ++it) {
        const int p = lane + 64 * it;
        if (TOTAL % 64 != 0 && p >= TOTAL) break;
        int x, y;
        if constexpr (KIND == P_YLINE) { x = (p / N) * S; y = (2 * (p % N) + 1) * U; }
        else if constexpr (KIND == P_XLINE) { y = (p / N) * (EXCL ? 1 : S); x = (2 * (p % N) + 1) * U; }
        else { y = (2 * (p / N) + 1) * U; x = (2 * (p % N) + 1) * U; }
        if (!(t.ox + x < t.g.nx && t.oy + y < t.g.ny)) continue;
        T pred;
        if constexpr (KIND == P_YLINE) pred = pred_1d<T, U, 1>(t, x, y, 0);
        else if constexpr (KIND == P_XLINE) pred = pred_1d<T, U, 0>(t, x, y, 0);
        else pred = pred_face<T, U>(t, x, y, 0);
        T& v = t.s[sidx(x, y, 0)];
        if constexpr (M == COMPRESS) {
            const int32_t qv = quantize(t, v, pred);
            if (x < 16 && y < 16) t.qc_out[lm_index<LEVEL>(t.g, t.ox + x, t.oy + y, 0)] = qv;
        } else {
            v = pred + (T)slot_code(v) * t.ebx2;
        }
    }
    wave_sync();
}

template <typename T, int U, int M>
__device__ __forceinline__ void plane_level(Ctx<T>& t, int lane, bool md, bool rev) {
    constexpr bool EX = U == 1;
    if (md) {  // LINE (x-lines and y-lines at even multiples), then FACE
        plane_stage<T, U, P_XLINE, 2 * U, false, M>(t, lane);
        plane_stage<T, U, P_YLINE, 2 * U, false, M>(t, lane);
        plane_stage<T, U, P_FACE, 2 * U, false, M>(t, lane);
    } else if (rev) {  // HOLLOW_R, YELLOW_R (BLUE_R is empty)
        plane_stage<T, U, P_XLINE, 2 * U, false, M>(t, lane);
        plane_stage<T, U, P_YLINE, U, false, M>(t, lane);
    } else {  // YELLOW, HOLLOW (BLUE is empty); HOLLOW of level 0 is the exclusive last stage
        plane_stage<T, U, P_YLINE, 2 * U, false, M>(t, lane);
        plane_stage<T, U, P_XLINE, U, EX, M>(t, lane);
    }
}

// level of plane point (x, y) of a tile: 4 for the corners (anchors)
__device__ __forceinline__ int plane_point_level(int x, int y) {
    const int m = x | y | 16;
    return __builtin_ctz(m) > 4 ? 4 : __builtin_ctz(m);
}

template <typename T, int M>
__device__ __forceinline__ void plane_tiles(const SplineArgs<T>& a, T* lds, int64_t unit) {
    const Geom& g = a.g;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t tile = unit * kPlaneWaves + wave;
    if (tile >= (int64_t)g.gx * g.gy) return;  // no workgroup barrier follows
    const int bx = (int)(tile % g.gx), by = (int)(tile / g.gx);
    Ctx<T> t;
    t.s = lds + wave * kPlaneVals;
    t.ox = bx * 16;
    t.oy = by * 16;
    t.oz = 0;
    t.c1x = t.ox + 16 < g.nx;
    t.c1y = t.oy + 16 < g.ny;
    t.c1z = false;
    t.g = g;
    t.qc_out = a.qc_out;
    t.qc_in = a.qc_in;
    t.q16 = a.q16;
    t.natural = false;
    t.acc = 0.0;
    uint8_t fl[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) fl[i] = a.flags8[i];
    const int maxlevel = (M == DECOMPRESS && a.d_maxlevel) ? *a.d_maxlevel : 4;
    constexpr int IT = (kPlaneVals + 63) / 64;
    if constexpr (M == COMPRESS) {
        T v[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int i = lane + 64 * it, x = i % kTS, y = i / kTS;
            v[it] = (T)0;
            if (i < kPlaneVals && t.ox + x < g.nx && t.oy + y < g.ny)
                v[it] = a.data[(int64_t)(t.ox + x) + (int64_t)(t.oy + y) * g.nx];
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int i = lane + 64 * it;
            if (i < kPlaneVals) t.s[i] = v[it];
        }
        wave_sync();
        if (lane == 0) a.anchors[bx + (int64_t)by * g.gx] = t.s[0];
    } else {
        // every slot: its code (levels up to maxlevel), its anchor, or 0
        int32_t q[IT];
        T anc[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int i = lane + 64 * it, x = i % kTS, y = i / kTS;
            q[it] = 0;
            anc[it] = (T)0;
            if (i < kPlaneVals && t.ox + x < g.nx && t.oy + y < g.ny) {
                const int lv = plane_point_level(x, y);
                const int gx = t.ox + x, gy = t.oy + y;
                if (lv == 4) {
                    if (a.anchors_in) anc[it] = a.anchors_in[(gx >> 4) + (int64_t)(gy >> 4) * g.gx];
                } else if (lv <= maxlevel) {
                    uint32_t idx;
                    switch (lv) {
                        case 0: idx = lm_index<0>(g, gx, gy, 0); break;
                        case 1: idx = lm_index<1>(g, gx, gy, 0); break;
                        case 2: idx = lm_index<2>(g, gx, gy, 0); break;
                        default: idx = lm_index<3>(g, gx, gy, 0); break;
                    }
                    q[it] = code_at(t.qc_in, t.q16, idx);
                }
            }
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int i = lane + 64 * it;
            if (i < kPlaneVals) t.s[i] = plane_point_level(i % kTS, i / kTS) == 4 ? anc[it] : code_slot(q[it], T());
        }
        wave_sync();
    }
    const LevelEb<T>& le = a.le;
    if (3 <= maxlevel) { t.eb_r = le.eb_r[3]; t.ebx2 = le.ebx2[3]; plane_level<T, 8, M>(t, lane, fl[3], fl[7]); }
    if (2 <= maxlevel) { t.eb_r = le.eb_r[2]; t.ebx2 = le.ebx2[2]; plane_level<T, 4, M>(t, lane, fl[2], fl[6]); }
    if (1 <= maxlevel) { t.eb_r = le.eb_r[1]; t.ebx2 = le.ebx2[1]; plane_level<T, 2, M>(t, lane, fl[1], fl[5]); }
    if (0 <= maxlevel) { t.eb_r = le.eb_r[0]; t.ebx2 = le.ebx2[0]; plane_level<T, 1, M>(t, lane, fl[0], fl[4]); }
    if constexpr (M == DECOMPRESS) {
#pragma unroll
        for (int it = 0; it < 4; ++it) {  // the 16 x 16 own points: lane -> x = lane mod 16, y = lane / 16 + 4 it
            const int x = lane & 15, y = (lane >> 4) + 4 * it;
            const int gx = t.ox + x, gy = t.oy + y;
            if (gx < g.nx && gy < g.ny) {
                const int64_t o = (int64_t)gx + (int64_t)gy * g.nx;
                const T v = t.s[sidx(x, y, 0)];
                __builtin_nontemporal_store(a.old ? v + a.old[o] : v, &a.out[o]);
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ SplineArgs<T> resolve(const SplineArgs<T>& a) {
    SplineArgs<T> r = a;
    if (a.le_dev) r.le = load_level_ebs<T>(a.le_dev);
    if (a.narrow) {
        const int32_t src = *a.narrow;
        r.q16 = src == 1;
        r.planes = src == 2 && a.ps.bp != nullptr;
    }
    return r;
}

template <typename T, int M>
__global__ __launch_bounds__(kThreads) void plane_kernel(SplineArgs<T> a) {
    __shared__ T lds[kPlaneWaves * kPlaneVals];
    clear_gaps<T, M>(a);
    plane_tiles<T, M>(resolve(a), lds, blockIdx.x);
}

__host__ __device__ inline bool plane_field(const Geom& g) { return g.nz == 1 && g.ny > 1; }

// ------------------------------------------------------------ 1-D fields --
// A field with ny = nz = 1 holds only the line y = z = 0 of each tile: no
// point of a y or z stage, and of the multi-dimensional stages only the
// LINE points along x, lies inside it, so every level reduces to the x
// stage of the line rule above (same stencil choice, same operations)
// whatever the flags say.  Lane l owns one whole tile, its 16 points and
// the halo point, in registers; a workgroup covers 256 consecutive tiles.
// Field values, codes and output pass through LDS so that every global
// access is coalesced.
constexpr int kLineTiles = kThreads;             // tiles per workgroup, one per lane
constexpr int kLinePts = 16 * kLineTiles;        // field points per workgroup
constexpr int kLineSlots = 17 * kLineTiles + 1;  // staged values: tile l point k at 17 l + k (k < 16)
constexpr int kLineCodes = 15 * kLineTiles;      // codes per workgroup (levels 3..0: 1, 2, 4, 8 per tile)

__device__ __forceinline__ int line_slot(int i) { return i + (i >> 4); }

template <int L>
constexpr int line_code_off() {  // level L's codes in the LDS code block
    return L == 3 ? 0 : (L == 2 ? kLineTiles : (L == 1 ? 3 * kLineTiles : 7 * kLineTiles));
}

// Points x = (2j + 1) U of one tile, j < N = 8 / U, in the order of the
// line stage: values at even multiples of U are known.
template <typename T, int L, int M>
__device__ __forceinline__ void line_level(T (&v)[17], int32_t (&q)[8 >> L], T eb_r, T ebx2, int ox, int nx,
                                           bool c1) {
    constexpr int U = 1 << L, N = 8 >> L;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int c = (2 * j + 1) * U;
        if (ox + c >= nx) break;  // beyond the field; so are later points
        const T pred = apply_stencil(stencil_of<U>(j, N, c1, ox, nx), v[j > 0 ? c - 3 * U : 0], v[c - U], v[c + U],
                                     v[j + 2 <= N ? c + 3 * U : 0]);
        if constexpr (M == COMPRESS) q[j] = quantize_val(v[c], pred, eb_r, ebx2);
        else v[c] = pred + (T)q[j] * ebx2;
    }
}

// codes of level L of the workgroup (code i = point m = unit 256 N + i of
// the level, at x = (2m + 1) U) between global memory and the LDS block
template <int L>
__device__ __forceinline__ void line_codes_out(const Geom& g, int64_t unit, const int32_t* ci, int32_t* qc) {
    constexpr int U = 1 << L, N = 8 >> L;
    const int lane = lane_id();
    const int64_t m0 = unit * (int64_t)(kLineTiles * N);
#pragma unroll
    for (int it = 0; it < N; ++it) {
        const int i = lane + it * kThreads;
        const int64_t m = m0 + i;
        if ((2 * m + 1) * U < g.nx) qc[g.prefix[L] + m] = ci[line_code_off<L>() + i];
    }
}

template <int L>
__device__ __forceinline__ void line_codes_load(const Geom& g, int64_t unit, bool load, const int32_t* qc, bool q16,
                                                int32_t (&r)[8 >> L]) {
    constexpr int U = 1 << L, N = 8 >> L;
    const int lane = lane_id();
    const int64_t m0 = unit * (int64_t)(kLineTiles * N);
#pragma unroll
    for (int it = 0; it < N; ++it) {
        const int64_t m = m0 + lane + it * kThreads;
        r[it] = (load && (2 * m + 1) * U < g.nx) ? code_at(qc, q16, (uint32_t)(g.prefix[L] + m)) : 0;
    }
}

template <int L>
__device__ __forceinline__ void line_codes_stage(const int32_t (&r)[8 >> L], int32_t* ci) {
    const int lane = lane_id();
#pragma unroll
    for (int it = 0; it < (8 >> L); ++it) ci[line_code_off<L>() + lane + it * kThreads] = r[it];
}

template <int L>
__device__ __forceinline__ void line_codes_take(const int32_t* ci, int32_t (&q)[8 >> L]) {
    constexpr int N = 8 >> L;
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) q[j] = ci[line_code_off<L>() + lane * N + j];
}

template <typename T, int M>
__device__ __forceinline__ void line_tiles(const SplineArgs<T>& a, T* lds, int64_t unit) {
    static_assert(kLineCodes * sizeof(int32_t) <= kLineSlots * sizeof(float), "code block fits the staging area");
    const Geom& g = a.g;
    const int nx = g.nx;
    const int64_t p0 = unit * kLinePts;
    const int lane = lane_id();
    const int64_t tile = unit * kLineTiles + lane;
    const bool active = tile < g.gx;
    const int ox = (int)(tile * 16);
    const bool c1 = (int64_t)ox + 16 < nx;
    int32_t* ci = reinterpret_cast<int32_t*>(lds);
    int32_t q3[1] = {0}, q2[2] = {0, 0}, q1[4] = {0, 0, 0, 0}, q0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    T v[17];
    if constexpr (M == COMPRESS) {
        constexpr int IT = (kLinePts + 1 + kThreads - 1) / kThreads;
        T s[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int i = lane + it * kThreads;
            s[it] = (i <= kLinePts && p0 + i < nx) ? a.data[p0 + i] : (T)0;
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int i = lane + it * kThreads;
            if (i <= kLinePts) lds[line_slot(i)] = s[it];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = lds[17 * lane + k];
        v[16] = lds[17 * (lane + 1)];  // halo: point 0 of the next tile
        if (active) a.anchors[tile] = v[0];
        __syncthreads();  // the staging area becomes the code block
        const LevelEb<T>& le = a.le;
        line_level<T, 3, M>(v, q3, le.eb_r[3], le.ebx2[3], ox, nx, c1);
        line_level<T, 2, M>(v, q2, le.eb_r[2], le.ebx2[2], ox, nx, c1);
        line_level<T, 1, M>(v, q1, le.eb_r[1], le.ebx2[1], ox, nx, c1);
        line_level<T, 0, M>(v, q0, le.eb_r[0], le.ebx2[0], ox, nx, c1);
        ci[line_code_off<3>() + lane] = q3[0];
#pragma unroll
        for (int j = 0; j < 2; ++j) ci[line_code_off<2>() + 2 * lane + j] = q2[j];
#pragma unroll
        for (int j = 0; j < 4; ++j) ci[line_code_off<1>() + 4 * lane + j] = q1[j];
#pragma unroll
        for (int j = 0; j < 8; ++j) ci[line_code_off<0>() + 8 * lane + j] = q0[j];
        __syncthreads();
        line_codes_out<3>(g, unit, ci, a.qc_out);
        line_codes_out<2>(g, unit, ci, a.qc_out);
        line_codes_out<1>(g, unit, ci, a.qc_out);
        line_codes_out<0>(g, unit, ci, a.qc_out);
    } else {
        const int maxlevel = a.d_maxlevel ? *a.d_maxlevel : 4;
        constexpr int OIT = kLinePts / kThreads;
        T prev[OIT];
        if (a.old) {
#pragma unroll
            for (int it = 0; it < OIT; ++it) {
                const int64_t x = p0 + lane + it * kThreads;
                prev[it] = x < nx ? a.old[x] : (T)0;
            }
        }
        int32_t r3[1], r2[2], r1[4], r0[8];
        line_codes_load<3>(g, unit, 3 <= maxlevel, a.qc_in, a.q16, r3);
        line_codes_load<2>(g, unit, 2 <= maxlevel, a.qc_in, a.q16, r2);
        line_codes_load<1>(g, unit, 1 <= maxlevel, a.qc_in, a.q16, r1);
        line_codes_load<0>(g, unit, 0 <= maxlevel, a.qc_in, a.q16, r0);
        const T a0 = (a.anchors_in && active) ? a.anchors_in[tile] : (T)0;
        const T a16 = (a.anchors_in && tile + 1 < g.gx) ? a.anchors_in[tile + 1] : (T)0;
        line_codes_stage<3>(r3, ci);
        line_codes_stage<2>(r2, ci);
        line_codes_stage<1>(r1, ci);
        line_codes_stage<0>(r0, ci);
        __syncthreads();
        line_codes_take<3>(ci, q3);
        line_codes_take<2>(ci, q2);
        line_codes_take<1>(ci, q1);
        line_codes_take<0>(ci, q0);
#pragma unroll
        for (int k = 1; k < 16; ++k) v[k] = (T)0;  // points of skipped levels stay 0
        v[0] = a0;
        v[16] = a16;
        const LevelEb<T>& le = a.le;
        if (3 <= maxlevel) line_level<T, 3, M>(v, q3, le.eb_r[3], le.ebx2[3], ox, nx, c1);
        if (2 <= maxlevel) line_level<T, 2, M>(v, q2, le.eb_r[2], le.ebx2[2], ox, nx, c1);
        if (1 <= maxlevel) line_level<T, 1, M>(v, q1, le.eb_r[1], le.ebx2[1], ox, nx, c1);
        if (0 <= maxlevel) line_level<T, 0, M>(v, q0, le.eb_r[0], le.ebx2[0], ox, nx, c1);
        __syncthreads();  // the code block becomes the output staging area
#pragma unroll
        for (int k = 0; k < 16; ++k) lds[17 * lane + k] = v[k];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < OIT; ++it) {
            const int i = lane + it * kThreads;
            if (p0 + i < nx) {
                const T r = lds[line_slot(i)];
                __builtin_nontemporal_store(a.old ? r + prev[it] : r, &a.out[p0 + i]);
            }
        }
    }
}

template <typename T, int M>
__global__ __launch_bounds__(kThreads) void line_kernel(SplineArgs<T> a) {
    __shared__ T lds[kLineSlots];
    clear_gaps<T, M>(a);
    line_tiles<T, M>(resolve(a), lds, blockIdx.x);
}

__host__ __device__ inline bool line_field(const Geom& g) { return g.ny == 1 && g.nz == 1; }

// kernel `interpolation` (compress) and `interpolation_reverse` /
// `interpolation_progressive_reverse` (decompress: anchors, or zero anchors
// plus old for the delta form).  Tiles away from the far faces of the field
// (all 17^3 points inside) run a variant whose boundary tests are resolved
// at compile time.
template <typename T, int M>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void spline_kernel(SplineArgs<T> a) {
    __shared__ T tile[kTileVals];
    __shared__ int32_t rowtab[M == DECOMPRESS ? kRowTab : 1];
    clear_gaps<T, M>(a);
    // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs,
    // so workgroup b runs on XCD b mod 8; each XCD takes a contiguous range
    // of tiles (x fastest) and the cache lines that neighbouring tiles share
    // (halo rows and columns, level-major code runs) meet in one L2.
    const Geom& g = a.g;
    const int64_t ntiles = (int64_t)g.gx * g.gy * g.gz;
    const int64_t b = blockIdx.x, xcd = b % kXcds, per = ntiles / kXcds, extra = ntiles % kXcds;
    const int64_t tix = xcd * per + (xcd < extra ? xcd : extra) + b / kXcds;
    const int bx = (int)(tix % g.gx), by = (int)((tix / g.gx) % g.gy), bz = (int)(tix / ((int64_t)g.gx * g.gy));
    const bool interior = bx * 16 + 16 < g.nx && by * 16 + 16 < g.ny && bz * 16 + 16 < g.nz;
    const SplineArgs<T> r = resolve(a);
    if constexpr (M == DECOMPRESS) {
        if (r.planes) {
            if (interior) tile_run_dec<T, true, CODES_PLANES>(r, tile, rowtab, bx, by, bz);
            else tile_run_dec<T, false, CODES_PLANES>(r, tile, rowtab, bx, by, bz);
            return;
        }
        if (r.q16) {
            if (interior) tile_run_dec<T, true, CODES_16>(r, tile, rowtab, bx, by, bz);
            else tile_run_dec<T, false, CODES_16>(r, tile, rowtab, bx, by, bz);
            return;
        }
        if (interior) tile_run_dec<T, true, CODES_32>(r, tile, rowtab, bx, by, bz);
        else tile_run_dec<T, false, CODES_32>(r, tile, rowtab, bx, by, bz);
    } else {
        if (interior) tile_run<T, M, true>(r, tile, rowtab, bx, by, bz);
        else tile_run<T, M, false>(r, tile, rowtab, bx, by, bz);
    }
}


// ---------------------------------------------------------------- tuner --
// Grid (sample block, candidate).  Lane p mod 256 of each stage accumulates
// |x - pred| in double in visiting order; lanes are added by a halving tree.
template <typename T>
__device__ __forceinline__ void tune_block(const T* __restrict__ data, const Geom& g, int sx0, int sy0, int sz0,
                                           int nsx, int nsy, int bid, int cand, double* __restrict__ part, T* tile,
                                           double* red) {
    Ctx<T> t;
    t.s = tile;
    t.ox = sx0 + (bid % nsx) * 128;
    t.oy = sy0 + ((bid / nsx) % nsy) * 128;
    t.oz = sz0 + (bid / nsx / nsy) * 128;
    t.c1x = t.ox + 16 < g.nx;
    t.c1y = t.oy + 16 < g.ny;
    t.c1z = t.oz + 16 < g.nz;
    t.g = g;
    t.qc_out = nullptr;
    t.qc_in = nullptr;
    t.q16 = false;
    t.acc = 0.0;
    load_tile(t, data);
    __syncthreads();
    // candidate table of auto_tuning_interpolation (pre_compute_att 110-141, 898-962)
    t.natural = cand >= 9 && cand <= 11;
    switch (cand) {
        case 0: level<T, 8, TUNE>(t, false, false); break;
        case 1: level<T, 8, TUNE>(t, false, true); break;
        case 2: level<T, 8, TUNE>(t, true, true); break;
        case 3: level<T, 4, TUNE>(t, false, false); break;
        case 4: level<T, 4, TUNE>(t, false, true); break;
        case 5: level<T, 4, TUNE>(t, true, true); break;
        case 6: case 9: level<T, 2, TUNE>(t, false, false); break;
        case 7: case 10: level<T, 2, TUNE>(t, false, true); break;
        case 8: case 11: level<T, 2, TUNE>(t, true, true); break;
        case 12: level<T, 1, TUNE>(t, false, false); break;
        case 13: level<T, 1, TUNE>(t, false, true); break;
        default: level<T, 1, TUNE>(t, true, true); break;
    }
    red[threadIdx.x] = t.acc;
    __syncthreads();
    for (int h = kThreads / 2; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)bid * 15 + cand] = red[0];
}

// Grid (sample block, candidate).  Lane p mod 256 of each stage accumulates
// |x - pred| in double in visiting order; lanes are added by a halving tree.
template <typename T>
__global__ __launch_bounds__(kThreads) void tune_kernel(const T* __restrict__ data, Geom g, int sx0,
                                                        int sy0, int sz0, int nsx, int nsy,
                                                        double* __restrict__ part) {
    __shared__ T tile[kTileVals];
    __shared__ double red[kThreads];
    tune_block<T>(data, g, sx0, sy0, sz0, nsx, nsy, blockIdx.x, blockIdx.y, part, tile, red);
}

// Sum over sample blocks in block order, then the host's pick rule
// (InterpolationPredictor.cu:2306-2355) on the device.
__device__ void pick_rule(const double* e, uint8_t* __restrict__ flags);

__device__ void pick_flags(const double* __restrict__ part, int nblocks, double* __restrict__ scores,
                           uint8_t* __restrict__ flags) {
    double e[15];
    for (int c = 0; c < 15; ++c) {
        double s = 0.0;
        for (int b = 0; b < nblocks; ++b) s += part[(int64_t)b * 15 + c];
        e[c] = s;
        if (scores) scores[c] = s;
    }
    pick_rule(e, flags);
}

// The pick rule on the 15 candidate sums.
__device__ void pick_rule(const double* e, uint8_t* __restrict__ flags) {
    uint8_t f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double best;
    if (e[0] > e[1]) { best = e[1]; f[7] = 1; } else { best = e[0]; f[7] = 0; }
    f[3] = e[2] < best;
    if (e[3] > e[4]) { best = e[4]; f[6] = 1; } else { best = e[3]; f[6] = 0; }
    f[2] = e[5] < best;
    best = e[6];
    int idx = 6;
    for (int i = 6; i < 12; i++) if (e[i] < best) { best = e[i]; idx = i; }
    f[1] = (idx == 8 || idx == 11);
    f[5] = (idx % 3) != 0;
    best = e[12];
    idx = 12;
    for (int i = 12; i < 15; i++) if (e[i] < best) { best = e[i]; idx = i; }
    f[0] = (idx == 14);
    f[4] = (idx % 3) != 0;
    for (int i = 0; i < 8; ++i) flags[i] = f[i];
}

// The partial sums go through LDS first (one cooperative load instead of a
// chain of dependent global loads in the one thread that adds them in order).
// Lane c < 15 then adds candidate c's sums in block order (the order of
// pick_flags), and lane 0 applies the rule.
constexpr int kPickLds = 2048;
__device__ __forceinline__ void pick_parallel(const double* __restrict__ part, int nblocks, double* __restrict__ scores,
                                              uint8_t* __restrict__ flags, double* s, double* e) {
    const int n = nblocks * 15;
    const bool staged = n <= kPickLds;
    if (staged)
        for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = part[i];
    __syncthreads();
    const double* src = staged ? s : part;
    if (threadIdx.x < 15) {
        const int c = threadIdx.x;
        double acc = 0.0;
        for (int b = 0; b < nblocks; ++b) acc += src[(int64_t)b * 15 + c];
        e[c] = acc;
        if (scores) scores[c] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) pick_rule(e, flags);
}

__global__ void tune_pick_kernel(const double* __restrict__ part, int nblocks, double* __restrict__ scores,
                                 uint8_t* __restrict__ flags) {
    __shared__ double s[kPickLds];
    __shared__ double e[15];
    pick_parallel(part, nblocks, scores, flags, s, e);
}

// ------------------------------------------------ extrema beside the tuner --
// Both only read the field, so one launch can run them side by side:
// workgroups [0, 15 nblocks) are the tuner's (sample block, candidate) pairs
// (the blocks of tune_kernel), the rest stream the extrema, `span` 16-byte
// vectors each (extrema_span).  A large field has few tuner blocks (135 for
// 384 x 384 x 256), too few to fill the GPU on their own.
template <typename T>
__global__ __launch_bounds__(kThreads) void extrema_tune_kernel(const T* __restrict__ data, int64_t n,
                                                                unsigned long long* __restrict__ keys, int64_t span,
                                                                Geom g, int sx0, int sy0, int sz0, int nsx, int nsy,
                                                                int nblocks, double* __restrict__ part) {
    __shared__ T tile[kTileVals];
    __shared__ double red[kThreads];
    const int ntune = 15 * nblocks;
    if ((int)blockIdx.x < ntune)
        tune_block<T>(data, g, sx0, sy0, sz0, nsx, nsy, (int)blockIdx.x % nblocks, (int)blockIdx.x / nblocks, part,
                      tile, red);
    else
        extrema_span<T>(data, n, (int64_t)blockIdx.x - ntune, span, keys);
}

__global__ void extrema_keys_init_kernel(unsigned long long* keys) {
    keys[0] = ~0ull;  // running min
    keys[1] = 0ull;   // running max
}

__device__ int64_t fill_bounds(const double* minmax, double eb_in, int rel, int dtype, double alpha, double* e);

// min / max from the keys, the tuner's pick (as tune_pick_kernel) and, with
// `bounds`, the level bounds of a relative eb (as level_bounds_kernel)
__global__ void extrema_tune_finish_kernel(const unsigned long long* __restrict__ keys, double* __restrict__ minmax,
                                           const double* __restrict__ part, int nblocks, double* __restrict__ scores,
                                           uint8_t* __restrict__ flags, double eb, int dtype, double alpha,
                                           double* __restrict__ bounds) {
    __shared__ double s[kPickLds];
    __shared__ double e[15];
    pick_parallel(part, nblocks, scores, flags, s, e);
    if (threadIdx.x != 0) return;
    minmax[0] = key_value(keys[0]);
    minmax[1] = key_value(keys[1]);
    if (bounds) {
        bounds[14] = (double)fill_bounds(minmax, eb, true, dtype, alpha, bounds);
        bounds[15] = 0.0;
    }
}

__global__ void zero_flags_kernel(double* scores, uint8_t* flags) {
    const int i = threadIdx.x;
    if (i < 8) flags[i] = 0;
    if (scores && i < 15) scores[i] = 0.0;
}

// ------------------------------------------------------- error bounds --
__host__ __device__ double alpha_of(double rel_eb) {  // InterpolationPredictor.cu:2248-2271
    const double a1 = 2.0, a2 = 1.75, a3 = 1.5, a4 = 1.25, a5 = 1;
    const double e1 = 1e-1, e2 = 1e-2, e3 = 1e-3, e4 = 1e-4, e5 = 1e-5;
    if (rel_eb >= e1) return a1;
    if (rel_eb >= e2) return a2 + (a1 - a2) * (rel_eb - e2) / (e1 - e2);
    if (rel_eb >= e3) return a3 + (a2 - a3) * (rel_eb - e3) / (e2 - e3);
    if (rel_eb >= e4) return a4 + (a3 - a4) * (rel_eb - e4) / (e3 - e4);
    if (rel_eb >= e5) return a5 + (a4 - a5) * (rel_eb - e5) / (e4 - e5);
    return a5;
}

// calc_eb (InterpolationPredictor.cu:1825-1839): every step rounds to T.
// Used on the host (plain launches) and on the device (batched launches,
// where the absolute bound comes from the device-side value range); the
// steps are single IEEE double operations, correctly rounded on both.
template <typename T>
__host__ __device__ LevelEb<T> level_ebs_alpha(double eb, double alpha) {
    LevelEb<T> le;
    const T ebx2 = (T)(eb * 2), eb_r = (T)(1 / eb);
    const double beta = 4.0;
    for (int lv = 0; lv < 4; ++lv) {
        const int unit = 1 << lv;
        T r = eb_r, x2 = ebx2;
        int temp = 1;
        while (temp < unit) {
            temp *= 2;
            r = (T)((double)r * alpha);
            x2 = (T)((double)x2 / alpha);
        }
        if ((double)x2 < (double)ebx2 / beta) {
            x2 = (T)((double)ebx2 / beta);
            r = (T)((double)eb_r * beta);
        }
        le.eb_r[lv] = r;
        le.ebx2[lv] = x2;
    }
    return le;
}

template <typename T>
LevelEb<T> level_ebs(double eb, double rel_eb) {
    return level_ebs_alpha<T>(eb, alpha_of(rel_eb));
}

// ------------------------------------------------------------- batched --
// Per field: the absolute bound from the value range (REL) and its level
// bounds; a bound that is not positive (a constant field in REL mode) marks
// the field failed and the stages run on a placeholder bound of 1.
// e = {absolute eb, range, eb_r[4], ebx2[4] (float64), eb_r[4] | ebx2[4]
// (float32)}; returns the status (PA_ERR_ARG for a bound that is not
// positive, the stages then running on a placeholder bound of 1)
__device__ int64_t fill_bounds(const double* minmax, double eb_in, int rel, int dtype, double alpha, double* e) {
    double range = 1.0, eb = eb_in;
    if (rel) {
        range = minmax[1] - minmax[0];
        eb = eb_in * range;
    }
    int64_t status = PA_OK;
    if (!(eb > 0)) {
        status = PA_ERR_ARG;
        eb = 1.0;
    }
    e[0] = eb;
    e[1] = range;
    if (dtype == PA_F32) {
        const LevelEb<float> le = level_ebs_alpha<float>(eb, alpha);
        float* q = reinterpret_cast<float*>(e + 10);
        for (int i = 0; i < 4; ++i) {
            q[i] = le.eb_r[i];
            q[4 + i] = le.ebx2[i];
        }
    } else {
        const LevelEb<double> le = level_ebs_alpha<double>(eb, alpha);
        for (int i = 0; i < 4; ++i) {
            e[2 + i] = le.eb_r[i];
            e[6 + i] = le.ebx2[i];
        }
    }
    return status;
}

__global__ void level_bounds_kernel(const double* minmax, double eb, int dtype, double alpha, double* bounds) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    bounds[14] = (double)fill_bounds(minmax, eb, minmax != nullptr, dtype, alpha, bounds);
    bounds[15] = 0.0;
}


// The batch's work before the bounds are known, in one launch: the tuner's
// (sample block, candidate) pairs of every field (each candidate's blocks
// one after the other, as the grid of tune_kernel), the extrema spans, the
// zeroing of the code arrays' gaps and the clearing of the encoder's
// look-back status words (one lane per launch slot of a chunk, then one per
// field for its error word).  None reads what another writes.  The tile is
// sized for a float64 tuner block; a float32 block uses half of it.
__global__ __launch_bounds__(kThreads) void batch_front_kernel(const BatchField* __restrict__ fields,
                                                               const int64_t* __restrict__ p_tune,
                                                               const int64_t* __restrict__ p_ext,
                                                               const int64_t* __restrict__ p_gap,
                                                               const int64_t* __restrict__ p_chunk, int nf,
                                                               int64_t n_tune, int64_t n_ext, int64_t n_gap,
                                                               int64_t n_chunk, int64_t span) {
    __shared__ double tile[kTileVals];
    __shared__ double red[kThreads];
    int64_t b = blockIdx.x;
    if (b < 15 * n_tune) {
        const int cand = (int)(b / n_tune);
        const int64_t u = b - cand * n_tune;
        const int f = find_field(p_tune, 0, nf, u);
        const BatchField& F = fields[f];
        const int bid = (int)(u - p_tune[f]);
        if (F.dtype == PA_F32)
            tune_block<float>(static_cast<const float*>(F.data), F.g, F.tsx0, F.tsy0, F.tsz0, F.tnsx, F.tnsy, bid,
                              cand, F.tune_part, reinterpret_cast<float*>(tile), red);
        else
            tune_block<double>(static_cast<const double*>(F.data), F.g, F.tsx0, F.tsy0, F.tsz0, F.tnsx, F.tnsy, bid,
                               cand, F.tune_part, tile, red);
        return;
    }
    b -= 15 * n_tune;
    if (b < n_ext) {
        const int f = find_field(p_ext, 0, nf, b);
        const BatchField& F = fields[f];
        unsigned long long* keys = reinterpret_cast<unsigned long long*>(F.minmax + 2);
        const int64_t wg = b - p_ext[f];
        if (F.dtype == PA_F32) extrema_span<float>(static_cast<const float*>(F.data), F.L.n, wg, span, keys);
        else extrema_span<double>(static_cast<const double*>(F.data), F.L.n, wg, span, keys);
        return;
    }
    b -= n_ext;
    if (b < n_gap) {
        const int f = find_field(p_gap, 0, nf, b);
        const BatchField& F = fields[f];
        zero_gaps(F.qc, F.gap_lo, F.gap_hi, b - p_gap[f], p_gap[f + 1] - p_gap[f]);
        return;
    }
    b -= n_gap;
    const int64_t i = b * kThreads + threadIdx.x;
    if (i >= n_chunk + nf) return;
    const int64_t gid = i < n_chunk ? i : 0;
    const int f = i < n_chunk ? find_field(p_chunk, 0, nf, gid) : (int)(i - n_chunk);
    const BatchField& F = fields[f];
    const int64_t k = i < n_chunk ? gid - p_chunk[f] : -1;  // status word k (-1: the error word)
    if (k < F.cg.nchunks) F.encws[1 + k] = 0ull;
}

// Per field, after the front launch: min and max from the extrema keys
// (then reset for the next run), the absolute bound and its level bounds,
// the tuner's pick (all flags false where the field has no tuner samples)
// and the field's result words.
__global__ __launch_bounds__(kThreads) void batch_finish_kernel(BatchField* fields) {
    __shared__ double s[kPickLds];
    __shared__ double e[15];
    BatchField& F = fields[blockIdx.x];
    if (F.tune_ok) {
        pick_parallel(F.tune_part, F.tnblocks, F.scores, F.flags, s, e);
    } else {
        if (threadIdx.x < 8) F.flags[threadIdx.x] = 0;
        if (threadIdx.x < 15) F.scores[threadIdx.x] = 0.0;
    }
    if (threadIdx.x != 0) return;
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(F.minmax + 2);
    F.minmax[0] = key_value(keys[0]);
    F.minmax[1] = key_value(keys[1]);
    keys[0] = ~0ull;
    keys[1] = 0ull;
    const int64_t status = fill_bounds(F.minmax, F.eb, F.rel, F.dtype, F.alpha, F.ebinfo);
    F.result[1] = (uint64_t)status;
    F.result[2] = (uint64_t)__double_as_longlong(F.ebinfo[1]);
    F.result[3] = (uint64_t)__double_as_longlong(F.ebinfo[0]);
}

// the extrema keys of every field at their start values (once per plan; the
// finish kernel resets them after each run)
__global__ void batch_keys_init_kernel(BatchField* fields, int nf) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(fields[f].minmax + 2);
    keys[0] = ~0ull;
    keys[1] = 0ull;
}

// Compression tiles of fields [lo, hi) (one dtype), with the XCD-aware
// order of spline_kernel taken over the concatenated list of work units: a
// tile, 4 tiles of a 2-D field or 256 tiles of a 1-D field.
template <typename T>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void batch_spline_kernel(
    const BatchField* __restrict__ fields, const int64_t* __restrict__ prefix, int lo, int hi, int64_t base,
    int64_t ntiles, int xcd_order) {
    __shared__ T tile[kTileVals];
    __shared__ int32_t rowtab[1];
    const int64_t b = blockIdx.x, xcd = b % kXcds, per = ntiles / kXcds, extra = ntiles % kXcds;
    const int64_t gt = base + (xcd_order ? xcd * per + (xcd < extra ? xcd : extra) + b / kXcds : b);
    const int f = find_field(prefix, lo, hi, gt);
    const BatchField& F = fields[f];
    const SplineArgs<T> a{static_cast<const T*>(F.data), F.g, load_level_ebs<T>(F.ebinfo), nullptr, nullptr, false,
                          F.flags, nullptr, F.qc, nullptr, static_cast<T*>(F.anchors), nullptr, nullptr, nullptr};
    const Geom& g = a.g;
    const int64_t tix = gt - prefix[f];
    if (line_field(g)) {
        line_tiles<T, COMPRESS>(a, tile, tix);
        return;
    }
    if (plane_field(g)) {
        plane_tiles<T, COMPRESS>(a, tile, tix);
        return;
    }
    const int bx = (int)(tix % g.gx), by = (int)((tix / g.gx) % g.gy), bz = (int)(tix / ((int64_t)g.gx * g.gy));
    const bool interior = bx * 16 + 16 < g.nx && by * 16 + 16 < g.ny && bz * 16 + 16 < g.nz;
    if (interior) tile_run<T, COMPRESS, true>(a, tile, rowtab, bx, by, bz);
    else tile_run<T, COMPRESS, false>(a, tile, rowtab, bx, by, bz);
}

// Reconstruction units of a batch of retrieval rounds, jobs [lo, hi) (one
// dtype): the anchored form on a field's first round, the delta form on old
// after that; levels above the round's maxlevel are skipped.
template <typename T>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void retr_spline_kernel(
    const RetrJob* __restrict__ jobs, const int64_t* __restrict__ prefix, int lo, int hi, int64_t base,
    int64_t nunits) {
    __shared__ T tile[kTileVals];
    __shared__ int32_t rowtab[kRowTab];
    const int64_t b = blockIdx.x, xcd = b % kXcds, per = nunits / kXcds, extra = nunits % kXcds;
    const int64_t gt = base + xcd * per + (xcd < extra ? xcd : extra) + b / kXcds;
    const int j = find_field(prefix, lo, hi, gt);
    const RetrJob& J = jobs[j];
    LevelEb<T> le;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        le.eb_r[i] = sizeof(T) == 4 ? (T)J.le32[i] : (T)J.le64[i];
        le.ebx2[i] = sizeof(T) == 4 ? (T)J.le32[4 + i] : (T)J.le64[4 + i];
    }
    // the round's code source: 0 int32 or 1 int16 codes in qc, 2 the planes
    const int32_t src = J.state[9];
    SplineArgs<T> a{nullptr, J.g, le, nullptr, nullptr, src == 1, J.flags, J.state + 8, nullptr, J.qc, nullptr,
                    J.first ? static_cast<const T*>(J.anchors) : nullptr,
                    J.first ? nullptr : static_cast<const T*>(J.old), static_cast<T*>(J.out)};
    if (src == 2) {
        a.ps.bp = J.bp;
        a.ps.begin = J.state + 10;  // each level's first live plane
        a.ps.end = J.state + 4;
        for (int lv = 0; lv < 4; ++lv) {
            a.ps.base[lv] = J.lv.base[lv];
            a.ps.stride[lv] = J.lv.stride[lv];
        }
        a.ps.size = J.lv.host_size;
        a.planes = true;
    }
    const Geom& g = a.g;
    const int64_t tix = gt - prefix[j];
    if (line_field(g)) {
        line_tiles<T, DECOMPRESS>(a, tile, tix);
        return;
    }
    if (plane_field(g)) {
        plane_tiles<T, DECOMPRESS>(a, tile, tix);
        return;
    }
    const int bx = (int)(tix % g.gx), by = (int)((tix / g.gx) % g.gy), bz = (int)(tix / ((int64_t)g.gx * g.gy));
    const bool interior = bx * 16 + 16 < g.nx && by * 16 + 16 < g.ny && bz * 16 + 16 < g.nz;
    if (a.planes) {
        if (interior) tile_run_dec<T, true, CODES_PLANES>(a, tile, rowtab, bx, by, bz);
        else tile_run_dec<T, false, CODES_PLANES>(a, tile, rowtab, bx, by, bz);
        return;
    }
    if (a.q16) {
        if (interior) tile_run_dec<T, true, CODES_16>(a, tile, rowtab, bx, by, bz);
        else tile_run_dec<T, false, CODES_16>(a, tile, rowtab, bx, by, bz);
        return;
    }
    if (interior) tile_run_dec<T, true, CODES_32>(a, tile, rowtab, bx, by, bz);
    else tile_run_dec<T, false, CODES_32>(a, tile, rowtab, bx, by, bz);
}

// ------------------------------------------------------------- host side --
Gaps make_gaps(const pa_layout* L) {
    Gaps gaps;
    const int64_t cnt[4] = {L->n - L->cnt[0], L->cnt[0] - L->cnt[1], L->cnt[1] - L->cnt[2], L->cnt[2] - L->n4};
    gaps.lo[0] = 0;
    gaps.hi[0] = L->n4;
    for (int lv = 3; lv >= 1; --lv) {
        gaps.lo[4 - lv] = L->prefix[lv] + cnt[lv];
        gaps.hi[4 - lv] = L->prefix[lv - 1];
    }
    gaps.lo[4] = L->real_end;
    gaps.hi[4] = L->code_len;
    return gaps;
}

int gap_blocks(const Gaps& gaps) {
    int64_t most = 0;
    for (int r = 0; r < 5; ++r) most = gaps.hi[r] - gaps.lo[r] > most ? gaps.hi[r] - gaps.lo[r] : most;
    return (int)((most + 255) / 256 < 1024 ? (most + 255) / 256 + 1 : 1024);
}

template <typename T, int M>
int launch_spline(const pa_layout* L, const T* data, double eb, double rel_eb, const uint8_t* flags8,
                  const int32_t* d_maxlevel, int32_t* qc_out, const int32_t* qc_in, T* anchors,
                  const T* anchors_in, const T* old, T* out, hipStream_t st, const double* le_dev = nullptr,
                  const int32_t* narrow = nullptr, const Gaps* gaps = nullptr, const PlaneSrc* ps = nullptr) {
    Geom g = make_geom(L);
    const int64_t ntiles = (int64_t)g.gx * g.gy * g.gz;
    if (ntiles > 0x7fffffff) return PA_ERR_ARG;
    LevelEb<T> le = le_dev ? LevelEb<T>{} : level_ebs<T>(eb, rel_eb);
    SplineArgs<T> args{data, g, le, le_dev, narrow, false, flags8, d_maxlevel, qc_out, qc_in, anchors, anchors_in,
                       old, out};
    if (gaps) {
        args.gaps = *gaps;
        args.gaps_on = 1;
    }
    if (ps) args.ps = *ps;
    if (line_field(g))
        hipLaunchKernelGGL((line_kernel<T, M>), dim3((unsigned)((ntiles + kLineTiles - 1) / kLineTiles)),
                           dim3(kThreads), 0, st, args);
    else if (plane_field(g))
        hipLaunchKernelGGL((plane_kernel<T, M>), dim3((unsigned)((ntiles + kPlaneWaves - 1) / kPlaneWaves)),
                           dim3(kThreads), 0, st, args);
    else
        hipLaunchKernelGGL((spline_kernel<T, M>), dim3((unsigned)ntiles), dim3(kThreads), 0, st, args);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

// calc_start_size with the reference's unsigned arithmetic (2277-2283):
// false where that arithmetic wraps (any axis shorter than 17).
bool tune_axis(uint32_t dim, int* start, int* size) {
    const uint32_t mid = dim / 2;
    if (mid < 8) return false;
    const int64_t tnum = (int64_t)dim - 8 - 1 - mid;
    if (tnum < 0) return false;
    const int k = (int)((mid - 8) / 128);
    *start = (int)mid - k * 128;
    *size = k + (int)(tnum / 128) + 1;
    return true;
}

bool valid_dtype(int d) { return d == PA_F32 || d == PA_F64; }

}  // namespace
}  // namespace pa

using namespace pa;

extern "C" size_t pa_autotune_workspace(const pa_layout* L) {
    int s[3], n[3];
    if (!L || !tune_axis(L->nx, &s[0], &n[0]) || !tune_axis(L->ny, &s[1], &n[1]) ||
        !tune_axis(L->nz, &s[2], &n[2]))
        return 0;
    return sizeof(double) * 15 * (size_t)n[0] * n[1] * n[2];
}

extern "C" int pa_autotune(const void* d_data, int dtype, const pa_layout* L, double* d_scores15,
                           uint8_t* d_flags8, void* d_ws, size_t ws_bytes, void* stream) {
    if (!L || !d_data || !d_flags8 || !valid_dtype(dtype)) return PA_ERR_ARG;
    hipStream_t st = as_stream(stream);
    int s[3], n[3];
    if (!tune_axis(L->nx, &s[0], &n[0]) || !tune_axis(L->ny, &s[1], &n[1]) || !tune_axis(L->nz, &s[2], &n[2])) {
        hipLaunchKernelGGL(zero_flags_kernel, dim3(1), dim3(64), 0, st, d_scores15, d_flags8);
        PA_LAUNCH_CHECK();
        return PA_OK;
    }
    const int nblocks = n[0] * n[1] * n[2];
    if (ws_bytes < pa_autotune_workspace(L) || !d_ws) return PA_ERR_WORKSPACE;
    double* part = static_cast<double*>(d_ws);
    Geom g = make_geom(L);
    dim3 grid(nblocks, 15);
    if (dtype == PA_F32)
        hipLaunchKernelGGL(tune_kernel<float>, grid, dim3(kThreads), 0, st, (const float*)d_data, g, s[0], s[1],
                           s[2], n[0], n[1], part);
    else
        hipLaunchKernelGGL(tune_kernel<double>, grid, dim3(kThreads), 0, st, (const double*)d_data, g, s[0],
                           s[1], s[2], n[0], n[1], part);
    PA_LAUNCH_CHECK();
    hipLaunchKernelGGL(tune_pick_kernel, dim3(1), dim3(64), 0, st, part, nblocks, d_scores15, d_flags8);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_extrema_autotune(const void* d_data, int dtype, const pa_layout* L, double* d_minmax, double rel_eb,
                                   double* d_bounds, double* d_scores15, uint8_t* d_flags8, void* d_ws,
                                   size_t ws_bytes, void* stream) {
    if (!L || !d_data || !d_minmax || !d_flags8 || !valid_dtype(dtype)) return PA_ERR_ARG;
    if (d_bounds && !(rel_eb > 0)) return PA_ERR_ARG;
    if (((uintptr_t)d_data & 15) != 0) return PA_ERR_ARG;
    int s[3], n[3];
    if (!tune_axis(L->nx, &s[0], &n[0]) || !tune_axis(L->ny, &s[1], &n[1]) || !tune_axis(L->nz, &s[2], &n[2])) {
        int e = pa_extrema(d_data, dtype, L->n, d_minmax, stream);  // no tuner samples: flags 0
        if (e == PA_OK) e = pa_autotune(d_data, dtype, L, d_scores15, d_flags8, d_ws, ws_bytes, stream);
        if (e == PA_OK && d_bounds) e = pa_level_bounds(d_minmax, rel_eb, dtype, d_bounds, stream);
        return e;
    }
    const int nblocks = n[0] * n[1] * n[2];
    if (ws_bytes < pa_autotune_workspace(L) || !d_ws) return PA_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    double* part = static_cast<double*>(d_ws);
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(d_minmax + 2);  // after the result, as pa_extrema
    hipLaunchKernelGGL(extrema_keys_init_kernel, dim3(1), dim3(1), 0, st, keys);
    PA_LAUNCH_CHECK();
    // the tuner's LDS tile leaves three workgroups per CU; the extrema are cut
    // into 8 workgroups per CU, queued behind the tuner blocks (measured
    // against 4 per CU and against exactly the slots the tuner leaves: the
    // step was 1 % faster)
    const int ntune = 15 * nblocks;
    const int64_t nv = L->n / (16 / dtype);
    int64_t ext = 8 * (int64_t)device_cus();
    int64_t span = (nv + ext - 1) / ext;
    if (span < 1) span = 1;
    ext = (nv + span - 1) / span;
    if (ext < 1) ext = 1;  // the last extrema workgroup also takes the scalar tail
    const unsigned grid = (unsigned)(ntune + ext);
    Geom g = make_geom(L);
    if (dtype == PA_F32)
        hipLaunchKernelGGL(extrema_tune_kernel<float>, dim3(grid), dim3(kThreads), 0, st, (const float*)d_data, L->n,
                           keys, span, g, s[0], s[1], s[2], n[0], n[1], nblocks, part);
    else
        hipLaunchKernelGGL(extrema_tune_kernel<double>, dim3(grid), dim3(kThreads), 0, st, (const double*)d_data,
                           L->n, keys, span, g, s[0], s[1], s[2], n[0], n[1], nblocks, part);
    PA_LAUNCH_CHECK();
    hipLaunchKernelGGL(extrema_tune_finish_kernel, dim3(1), dim3(64), 0, st, keys, d_minmax, part, nblocks, d_scores15,
                       d_flags8, rel_eb, dtype, alpha_of(rel_eb), d_bounds);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_spline_construct(const void* d_data, int dtype, const pa_layout* L, double eb,
                                   double rel_eb, const uint8_t* d_flags8, int32_t* d_qc, void* d_anchors,
                                   void* stream) {
    if (!L || !d_data || !d_flags8 || !d_qc || !d_anchors || !valid_dtype(dtype) || !(eb > 0))
        return PA_ERR_ARG;
    hipStream_t st = as_stream(stream);
    const Gaps gaps = make_gaps(L);
    if (dtype == PA_F32)
        return launch_spline<float, COMPRESS>(L, (const float*)d_data, eb, rel_eb, d_flags8, nullptr, d_qc,
                                              nullptr, (float*)d_anchors, nullptr, nullptr, nullptr, st, nullptr,
                                              nullptr, &gaps);
    return launch_spline<double, COMPRESS>(L, (const double*)d_data, eb, rel_eb, d_flags8, nullptr, d_qc,
                                           nullptr, (double*)d_anchors, nullptr, nullptr, nullptr, st, nullptr,
                                           nullptr, &gaps);
}

extern "C" int pa_level_bounds(const double* d_minmax, double eb, int dtype, double* d_bounds, void* stream) {
    if (!d_bounds || !valid_dtype(dtype) || !(eb > 0)) return PA_ERR_ARG;
    const double alpha = alpha_of(d_minmax ? eb : 0.0);
    hipLaunchKernelGGL(level_bounds_kernel, dim3(1), dim3(64), 0, as_stream(stream), d_minmax, eb, dtype, alpha,
                       d_bounds);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_spline_construct_dev(const void* d_data, int dtype, const pa_layout* L, const double* d_bounds,
                                       const uint8_t* d_flags8, int32_t* d_qc, void* d_anchors, void* stream) {
    if (!L || !d_data || !d_bounds || !d_flags8 || !d_qc || !d_anchors || !valid_dtype(dtype)) return PA_ERR_ARG;
    hipStream_t st = as_stream(stream);
    const Gaps gaps = make_gaps(L);
    if (dtype == PA_F32)
        return launch_spline<float, COMPRESS>(L, (const float*)d_data, 0.0, 0.0, d_flags8, nullptr, d_qc, nullptr,
                                              (float*)d_anchors, nullptr, nullptr, nullptr, st, d_bounds, nullptr,
                                              &gaps);
    return launch_spline<double, COMPRESS>(L, (const double*)d_data, 0.0, 0.0, d_flags8, nullptr, d_qc, nullptr,
                                           (double*)d_anchors, nullptr, nullptr, nullptr, st, d_bounds, nullptr,
                                           &gaps);
}

extern "C" int pa_spline_reconstruct(const int32_t* d_qc, const void* d_anchors, int dtype, const pa_layout* L,
                                     double eb, double rel_eb, const uint8_t* d_flags8,
                                     const int32_t* d_maxlevel, void* d_out, void* stream) {
    if (!L || !d_qc || !d_anchors || !d_flags8 || !d_out || !valid_dtype(dtype) || !(eb > 0)) return PA_ERR_ARG;
    hipStream_t st = as_stream(stream);
    if (dtype == PA_F32)
        return launch_spline<float, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, d_qc,
                                                nullptr, (const float*)d_anchors, nullptr, (float*)d_out, st);
    return launch_spline<double, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, d_qc, nullptr,
                                             (const double*)d_anchors, nullptr, (double*)d_out, st);
}

extern "C" int pa_spline_reconstruct_narrow(const void* d_qc, const int32_t* d_narrow, const void* d_anchors, int dtype,
                                            const pa_layout* L, double eb, double rel_eb, const uint8_t* d_flags8,
                                            const int32_t* d_maxlevel, void* d_out, void* stream) {
    if (!L || !d_qc || !d_narrow || !d_anchors || !d_flags8 || !d_out || !valid_dtype(dtype) || !(eb > 0))
        return PA_ERR_ARG;
    hipStream_t st = as_stream(stream);
    const int32_t* q = static_cast<const int32_t*>(d_qc);
    if (dtype == PA_F32)
        return launch_spline<float, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, q, nullptr,
                                                (const float*)d_anchors, nullptr, (float*)d_out, st, nullptr, d_narrow);
    return launch_spline<double, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, q, nullptr,
                                             (const double*)d_anchors, nullptr, (double*)d_out, st, nullptr, d_narrow);
}

extern "C" int pa_spline_delta_reconstruct_narrow(const void* d_qc, const int32_t* d_narrow, int dtype,
                                                  const pa_layout* L, double eb, double rel_eb,
                                                  const uint8_t* d_flags8, const int32_t* d_maxlevel,
                                                  const void* d_old, void* d_out, void* stream) {
    if (!L || !d_qc || !d_narrow || !d_flags8 || !d_old || !d_out || !valid_dtype(dtype) || !(eb > 0))
        return PA_ERR_ARG;
    hipStream_t st = as_stream(stream);
    const int32_t* q = static_cast<const int32_t*>(d_qc);
    if (dtype == PA_F32)
        return launch_spline<float, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, q, nullptr,
                                                nullptr, (const float*)d_old, (float*)d_out, st, nullptr, d_narrow);
    return launch_spline<double, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, q, nullptr,
                                             nullptr, (const double*)d_old, (double*)d_out, st, nullptr, d_narrow);
}

extern "C" int pa_spline_reconstruct_planes(const void* d_qc, const int32_t* d_src, const uint8_t* d_bp,
                                            const int32_t* d_begin, const int32_t* d_end, const void* d_anchors,
                                            const void* d_old, int dtype, const pa_layout* L, double eb,
                                            double rel_eb, const uint8_t* d_flags8, const int32_t* d_maxlevel,
                                            void* d_out, void* stream) {
    if (!L || !d_qc || !d_src || !d_bp || !d_begin || !d_end || !d_flags8 || !d_out || !valid_dtype(dtype) ||
        !(eb > 0) || (d_anchors == nullptr) == (d_old == nullptr))
        return PA_ERR_ARG;
    PlaneSrc ps{};
    ps.bp = d_bp;
    ps.begin = d_begin;
    ps.end = d_end;
    for (int lv = 0; lv < 4; ++lv) {
        ps.base[lv] = L->dev_base[lv];
        ps.stride[lv] = L->dev_stride[lv];
    }
    ps.size = L->host_size;
    hipStream_t st = as_stream(stream);
    const int32_t* q = static_cast<const int32_t*>(d_qc);
    if (dtype == PA_F32)
        return launch_spline<float, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, q, nullptr,
                                                (const float*)d_anchors, (const float*)d_old, (float*)d_out, st,
                                                nullptr, d_src, nullptr, &ps);
    return launch_spline<double, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, q, nullptr,
                                             (const double*)d_anchors, (const double*)d_old, (double*)d_out, st,
                                             nullptr, d_src, nullptr, &ps);
}

extern "C" int pa_spline_delta_reconstruct(const int32_t* d_qc, int dtype, const pa_layout* L, double eb,
                                           double rel_eb, const uint8_t* d_flags8, const int32_t* d_maxlevel,
                                           const void* d_old, void* d_out, void* stream) {
    if (!L || !d_qc || !d_flags8 || !d_old || !d_out || !valid_dtype(dtype) || !(eb > 0)) return PA_ERR_ARG;
    hipStream_t st = as_stream(stream);
    if (dtype == PA_F32)
        return launch_spline<float, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, d_qc,
                                                nullptr, nullptr, (const float*)d_old, (float*)d_out, st);
    return launch_spline<double, DECOMPRESS>(L, nullptr, eb, rel_eb, d_flags8, d_maxlevel, nullptr, d_qc, nullptr,
                                             nullptr, (const double*)d_old, (double*)d_out, st);
}

// ------------------------------------------------------------- batched --
int64_t pa::batch_gap_blocks(const BatchField& f) {
    Gaps g;
    for (int r = 0; r < 5; ++r) {
        g.lo[r] = f.gap_lo[r];
        g.hi[r] = f.gap_hi[r];
    }
    return gap_blocks(g);
}

void pa::batch_predict_setup(BatchField& f) {
    f.g = make_geom(&f.L);
    const Gaps g = make_gaps(&f.L);
    for (int r = 0; r < 5; ++r) {
        f.gap_lo[r] = g.lo[r];
        f.gap_hi[r] = g.hi[r];
    }
    f.alpha = alpha_of(f.rel ? f.eb : 0.0);
    int s[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    f.tune_ok = tune_axis(f.L.nx, &s[0], &n[0]) && tune_axis(f.L.ny, &s[1], &n[1]) && tune_axis(f.L.nz, &s[2], &n[2]);
    f.tsx0 = s[0];
    f.tsy0 = s[1];
    f.tsz0 = s[2];
    f.tnsx = n[0];
    f.tnsy = n[1];
    f.tnblocks = f.tune_ok ? n[0] * n[1] * n[2] : 0;
}

int pa::batch_front(const BatchDev& b, hipStream_t st) {
    const int64_t clr = (b.n_chunk + b.nf + kThreads - 1) / kThreads;
    const int64_t blocks = 15 * b.n_tune + b.n_ext + b.n_gap + clr;
    hipLaunchKernelGGL(batch_front_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, st, b.fields, b.p_tune, b.p_ext,
                       b.p_gap, b.p_chunk, b.nf, b.n_tune, b.n_ext, b.n_gap, b.n_chunk,
                       (int64_t)b.ext_batches * kExtBatch);
    PA_LAUNCH_CHECK();
    hipLaunchKernelGGL(batch_finish_kernel, dim3((unsigned)b.nf), dim3(kThreads), 0, st, b.fields);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

int pa::batch_keys_init(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL(batch_keys_init_kernel, dim3((unsigned)((b.nf + 255) / 256)), dim3(256), 0, st, b.fields, b.nf);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

int pa::batch_spline(const BatchDev& b, hipStream_t st) {
    if (b.n_tile32 > 0) {
        hipLaunchKernelGGL(batch_spline_kernel<float>, dim3((unsigned)b.n_tile32), dim3(kThreads), 0, st, b.fields,
                           b.p_tile, 0, b.nf32, (int64_t)0, b.n_tile32, b.xcd_order);
        PA_LAUNCH_CHECK();
    }
    if (b.n_tile - b.n_tile32 > 0) {
        hipLaunchKernelGGL(batch_spline_kernel<double>, dim3((unsigned)(b.n_tile - b.n_tile32)), dim3(kThreads), 0, st,
                           b.fields, b.p_tile, b.nf32, b.nf, b.n_tile32, b.n_tile - b.n_tile32, b.xcd_order);
        PA_LAUNCH_CHECK();
    }
    return PA_OK;
}

int64_t pa::batch_tile_units(const pa_layout* L) {
    const Geom g = make_geom(L);
    const int64_t ntiles = (int64_t)g.gx * g.gy * g.gz;
    if (line_field(g)) return (ntiles + kLineTiles - 1) / kLineTiles;
    if (plane_field(g)) return (ntiles + kPlaneWaves - 1) / kPlaneWaves;
    return ntiles;
}

void pa::retr_level_ebs(RetrJob& j, double eb, double rel_eb) {
    const LevelEb<float> l32 = level_ebs<float>(eb, rel_eb);
    const LevelEb<double> l64 = level_ebs<double>(eb, rel_eb);
    for (int i = 0; i < 4; ++i) {
        j.le32[i] = l32.eb_r[i];
        j.le32[4 + i] = l32.ebx2[i];
        j.le64[i] = l64.eb_r[i];
        j.le64[4 + i] = l64.ebx2[i];
    }
}

int pa::retr_reconstruct(const RetrDev& r, hipStream_t st) {
    if (r.n_unit32 > 0) {
        hipLaunchKernelGGL(retr_spline_kernel<float>, dim3((unsigned)r.n_unit32), dim3(kThreads), 0, st, r.jobs,
                           r.p_unit, 0, r.nj32, (int64_t)0, r.n_unit32);
        PA_LAUNCH_CHECK();
    }
    if (r.n_unit - r.n_unit32 > 0) {
        hipLaunchKernelGGL(retr_spline_kernel<double>, dim3((unsigned)(r.n_unit - r.n_unit32)), dim3(kThreads), 0, st,
                           r.jobs, r.p_unit, r.nj32, r.nj, r.n_unit32, r.n_unit - r.n_unit32);
        PA_LAUNCH_CHECK();
    }
    return PA_OK;
}
