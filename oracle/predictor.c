/*
 * predictor.c -- instantiates the oracle predictor for float and double
 * (test infrastructure only; see prism_oracle.h).
 */
#include "prism_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Host flag pick, InterpolationPredictor.cu:2306-2355 (strict comparisons,
 * first index wins ties). */
void po_pick_flags(const double e[15], uint8_t flags[8]) {
    double best;
    int idx;
    memset(flags, 0, 8);
    if (e[0] > e[1]) { best = e[1]; flags[4 + 3] = 1; }
    else { best = e[0]; flags[4 + 3] = 0; }
    flags[3] = e[2] < best;
    if (e[3] > e[4]) { best = e[4]; flags[4 + 2] = 1; }
    else { best = e[3]; flags[4 + 2] = 0; }
    flags[2] = e[5] < best;
    best = e[6];
    idx = 6;
    for (int i = 6; i < 12; i++)
        if (e[i] < best) { best = e[i]; idx = i; }
    flags[1] = (idx == 8 || idx == 11);
    flags[4 + 1] = (idx % 3) != 0;
    best = e[12];
    idx = 12;
    for (int i = 12; i < 15; i++)
        if (e[i] < best) { best = e[i]; idx = i; }
    flags[0] = (idx == 14);
    flags[4 + 0] = (idx % 3) != 0;
}

/* get_interp_order, InterpolationPredictor.cu:1538-1545 (c1 written as
 * "the tile is not the last one on this axis": origin + 16 < N, which is
 * the same test as BI != GD-1 for aligned tiles and the tuner's test at 286) */
static inline int interp_order(int c, int unit, int64_t g, int64_t N, int c1) {
    int b = (c >= 3 * unit) ? 3 : 1;
    int f = ((c + 3 * unit <= 16) && (c1 || (g + 3 * unit < N))) ? 3 : ((c1 || (g + unit < N)) ? 1 : 0);
    return (b == 3) ? ((f == 3) ? 4 : ((f == 1) ? 3 : 0)) : ((f == 3) ? 2 : ((f == 1) ? 1 : 0));
}

/* stage kinds */
enum { ST_BLUE, ST_YELLOW, ST_HOLLOW, ST_HOLLOW_R, ST_YELLOW_R, ST_BLUE_R, ST_LINE, ST_FACE, ST_CUBE };

/* calc_start_size with the reference's unsigned arithmetic, 2277-2283.
 * Returns -1 where the reference's unsigned subtraction wraps. */
static int tune_axis(uint32_t dim, int* start, int* size) {
    uint32_t mid = dim / 2;
    if (mid < 8) return -1;
    int64_t tnum = (int64_t)dim - 8 - 1 - mid;
    if (tnum < 0) return -1;
    int k = (int)((mid - 8) / 128);
    int tt = (int)(tnum / 128);
    *start = (int)mid - k * 128;
    *size = k + tt + 1;
    return 0;
}

/* Candidate table of auto_tuning_interpolation (pre_compute_att 110-141 and
 * 898-962): level_id, use_md, reverse, natural for scores[0..14]. */
static const int po_tune_cand[15][4] = {
    {3, 0, 0, 0}, {3, 0, 1, 0}, {3, 1, 1, 0}, {2, 0, 0, 0}, {2, 0, 1, 0},
    {2, 1, 1, 0}, {1, 0, 0, 0}, {1, 0, 1, 0}, {1, 1, 1, 0}, {1, 0, 0, 1},
    {1, 0, 1, 1}, {1, 1, 1, 1}, {0, 0, 0, 0}, {0, 0, 1, 0}, {0, 1, 1, 0}};

#define PO_T float
#define PO_SFX(name) name##_f32
#define PO_FABS fabsf
#define PO_LEVEL_EB po_level_eb_f32
#include "predictor_impl.inc"
#undef PO_T
#undef PO_SFX
#undef PO_FABS
#undef PO_LEVEL_EB
#undef TS
#undef SIDX

#define PO_T double
#define PO_SFX(name) name##_f64
#define PO_FABS fabs
#define PO_LEVEL_EB po_level_eb_f64
#include "predictor_impl.inc"
