/*
 * layout.c -- field geometry, level-major indexing, error-bound schedule and
 * extrema for the CPU oracle (test infrastructure only; see prism_oracle.h).
 */
#include "prism_oracle.h"

#include <math.h>
#include <string.h>

/* roundup of (x/8) to a multiple of 8, with the reference's quirk that an
 * empty segment still gets 8 bytes: (((x)/8 - 1)/8 + 1)*8 in C int
 * arithmetic (bitplane.cu:42-45). */
static int64_t dev_seg_stride(int64_t codes) {
    int64_t g = codes / 8;
    if (g == 0) return 8;
    return ((g - 1) / 8 + 1) * 8;
}

void po_layout_init(po_layout* L, uint32_t nx, uint32_t ny, uint32_t nz) {
    memset(L, 0, sizeof(*L));
    L->nx = nx;
    L->ny = ny;
    L->nz = nz;
    L->n = (int64_t)nx * ny * nz;

    /* level grids: InterpolationPredictor.cu:969-985 */
    int64_t dx = nx, dy = ny, dz = nz;
    for (int level = 0; level <= 4; ++level) {
        L->lx[level] = dx;
        L->lxy[level] = dx * dy;
        if (level < 4) {
            dx = (dx + 1) >> 1;
            dy = (dy + 1) >> 1;
            dz = (dz + 1) >> 1;
            L->cnt[level] = dx * dy * dz;
        }
    }
    L->n4 = L->cnt[3];

    /* device prefix with 8-alignment of every level region:
     * InterpolationPredictor.cu:986-992 (absolute) == bitplane.cu:22-35 (relative) */
    int64_t p[5];
    for (int i = 0; i < 4; ++i) p[i] = L->cnt[i];
    p[4] = 0;
    int64_t align = 0;
    for (int i = 2; i >= 0; --i) {
        align += (8 - ((p[i] - p[i + 1] + align) % 8)) % 8;
        p[i] += align;
    }
    int64_t pads_before_l0 = align;
    /* device tail pad counts the n4 phantom codes (bitplane.cu:35) */
    align += (8 - ((L->n - (p[0] - L->n4) + align) % 8)) % 8;
    for (int i = 0; i < 5; ++i) L->prefix[i] = p[i];
    L->dev_align = align;
    L->real_end = L->n + pads_before_l0;
    L->code_len = L->n4 + L->n + L->dev_align;

    /* relative level segments (codes) and device strides, bitplane.cu:42-49 */
    int64_t rp0 = p[0] - L->n4, rp1 = p[1] - L->n4, rp2 = p[2] - L->n4;
    L->dev_len[3] = rp2;
    L->dev_len[2] = rp1 - rp2;
    L->dev_len[1] = rp0 - rp1;
    L->dev_len[0] = L->n - rp0 + L->dev_align;
    for (int l = 0; l < 4; ++l) L->dev_stride[l] = dev_seg_stride(L->dev_len[l]);
    L->dev_base[3] = 0;
    L->dev_base[2] = 32 * L->dev_stride[3];
    L->dev_base[1] = L->dev_base[2] + 32 * L->dev_stride[2];
    L->dev_base[0] = L->dev_base[1] + 32 * L->dev_stride[1];
    L->dev_size = L->dev_base[0] + 32 * L->dev_stride[0];

    /* host strides: io.hpp:104-132 pads with 8-(r%8) (1..8, never 0),
     * then io.cu:292-302 rounds each to a multiple of 8 */
    int64_t h[5];
    for (int i = 0; i < 4; ++i) h[i] = L->cnt[i] - L->n4;
    int64_t halign = 0;
    for (int i = 2; i >= 0; --i) {
        halign += 8 - ((h[i] - h[i + 1] + halign) % 8);
        h[i] += halign;
    }
    halign += 8 - ((L->n - h[0] + halign) % 8);
    int64_t hs[4];
    hs[3] = (h[2] - h[3]) >> 3;
    hs[2] = (h[1] - h[2]) >> 3;
    hs[1] = (h[0] - h[1]) >> 3;
    hs[0] = (L->n - h[0] + halign) >> 3;
    for (int l = 0; l < 4; ++l) L->host_stride[l] = ((hs[l] + 7) / 8) * 8;
    L->host_base[3] = 0;
    L->host_base[2] = 32 * L->host_stride[3];
    L->host_base[1] = L->host_base[2] + 32 * L->host_stride[2];
    L->host_base[0] = L->host_base[1] + 32 * L->host_stride[1];
    L->host_size = L->host_base[0] + 32 * L->host_stride[0];

    /* codec chunk grid, zbpc.cu:176-178 */
    L->nchunks = 0;
    for (int l = 0; l < 4; ++l) {
        L->chunks_per_plane[l] = (L->host_stride[l] + PO_CS - 1) / PO_CS;
        L->nchunks += 32 * L->chunks_per_plane[l];
    }
    L->trailer_fixed = 4 + 4 + 8 + 4 * 32 * 8; /* compressor.cu:113 minus anchors */
}

int64_t po_lm_index(const po_layout* L, uint32_t gx, uint32_t gy, uint32_t gz) {
    int level = 0;
    while (level < 4 && (gx % 2 == 0) && (gy % 2 == 0) && (gz % 2 == 0)) {
        gx >>= 1;
        gy >>= 1;
        gz >>= 1;
        level++;
    }
    int64_t gid = (int64_t)gx + (int64_t)gy * L->lx[level] + (int64_t)gz * L->lxy[level];
    if (level < 4) {
        gid += L->prefix[level] - (int64_t)((gz + 1) >> 1) * L->lxy[level + 1] -
               (int64_t)(gz % 2 == 0) * (int64_t)((gy + 1) >> 1) * L->lx[level + 1] -
               (int64_t)(gz % 2 == 0 && gy % 2 == 0) * (int64_t)((gx + 1) >> 1);
    }
    return gid;
}

/* InterpolationPredictor.cu:2248-2271 */
double po_alpha(double rel_eb) {
    const double a1 = 2.0, a2 = 1.75, a3 = 1.5, a4 = 1.25, a5 = 1;
    const double e1 = 1e-1, e2 = 1e-2, e3 = 1e-3, e4 = 1e-4, e5 = 1e-5;
    if (rel_eb >= e1) return a1;
    if (rel_eb >= e2) return a2 + (a1 - a2) * (rel_eb - e2) / (e1 - e2);
    if (rel_eb >= e3) return a3 + (a2 - a3) * (rel_eb - e3) / (e2 - e3);
    if (rel_eb >= e4) return a4 + (a3 - a4) * (rel_eb - e4) / (e3 - e4);
    if (rel_eb >= e5) return a5 + (a4 - a5) * (rel_eb - e5) / (e4 - e5);
    return a5;
}

/* calc_eb, InterpolationPredictor.cu:1825-1839: every step rounds to T */
void po_level_eb_f32(float eb_r, float ebx2, double alpha, double beta, int unit, float* cur_eb_r,
                     float* cur_ebx2) {
    float r = eb_r, x2 = ebx2;
    int temp = 1;
    while (temp < unit) {
        temp *= 2;
        r = (float)((double)r * alpha);
        x2 = (float)((double)x2 / alpha);
    }
    if ((double)x2 < (double)ebx2 / beta) {
        x2 = (float)((double)ebx2 / beta);
        r = (float)((double)eb_r * beta);
    }
    *cur_eb_r = r;
    *cur_ebx2 = x2;
}

void po_level_eb_f64(double eb_r, double ebx2, double alpha, double beta, int unit,
                     double* cur_eb_r, double* cur_ebx2) {
    double r = eb_r, x2 = ebx2;
    int temp = 1;
    while (temp < unit) {
        temp *= 2;
        r *= alpha;
        x2 /= alpha;
    }
    if (x2 < ebx2 / beta) {
        x2 = ebx2 / beta;
        r = eb_r * beta;
    }
    *cur_eb_r = r;
    *cur_ebx2 = x2;
}

/* compressor.cu:228-302: exact min / max of the field */
void po_extrema_f32(const float* in, int64_t n, double out[2]) {
    float mn = in[0], mx = in[0];
    for (int64_t i = 1; i < n; ++i) {
        if (in[i] < mn) mn = in[i];
        if (in[i] > mx) mx = in[i];
    }
    out[0] = mn;
    out[1] = mx;
}

void po_extrema_f64(const double* in, int64_t n, double out[2]) {
    double mn = in[0], mx = in[0];
    for (int64_t i = 1; i < n; ++i) {
        if (in[i] < mn) mn = in[i];
        if (in[i] > mx) mx = in[i];
    }
    out[0] = mn;
    out[1] = mx;
}
