/*
 * bitplane.c -- bit-plane shuffle of the CPU oracle (test infrastructure
 * only; see prism_oracle.h).
 *
 * Closed form of bitplane.cu:55-253: for level l, plane k (0 = most
 * significant bit) and 8-code group g of that level's segment,
 *     bp[A_l + k*s_l + g] = sum_j bit_{31-k}(w[8g+j]) << (7-j)
 * where w = NB(code) or SM(code).  The test suite checks this closed form
 * against an emulation of the reference's shuffle/byte_perm/delta-swap
 * network (bitplane.cu:90-132).  Codes past the last real code of level 0
 * (the n4 phantom tail the reference reads past the code array,
 * bitplane.cu:178-186) are taken as 0.
 */
#include "prism_oracle.h"

#include <string.h>

#define NB_MASK 0xAAAAAAAAu

static inline uint32_t fwd(int32_t code, int mode) {
    uint32_t a = (uint32_t)code;
    if (mode == 0) return (a + NB_MASK) ^ NB_MASK;          /* bitplane.cu:80-81 */
    uint32_t sign = (uint32_t)(code >> 31);                   /* bitplane.cu:82-85 */
    return (a & 0x7FFFFFFFu) ^ sign;
}

static inline int32_t inv(uint32_t a, int mode) {
    if (mode == 0) return (int32_t)((a ^ NB_MASK) - NB_MASK); /* bitplane.cu:365-366 */
    uint32_t sign = (uint32_t)(((int32_t)a) >> 31);           /* bitplane.cu:367-370 */
    return (int32_t)((a & 0x7FFFFFFFu) ^ sign);
}

/* start (relative to the anchors) and real code count of each level segment */
static void level_segments(const po_layout* L, int64_t rel_start[4], int64_t real_cnt[4]) {
    rel_start[3] = 0;
    rel_start[2] = L->prefix[2] - L->n4;
    rel_start[1] = L->prefix[1] - L->n4;
    rel_start[0] = L->prefix[0] - L->n4;
    real_cnt[3] = L->cnt[2] - L->n4;
    real_cnt[2] = L->cnt[1] - L->cnt[2];
    real_cnt[1] = L->cnt[0] - L->cnt[1];
    real_cnt[0] = L->n - L->cnt[0];
}

void po_bitshuffle(const int32_t* qc, const po_layout* L, int mode, uint8_t* bp) {
    int64_t rs[4], rc[4];
    level_segments(L, rs, rc);
    memset(bp, 0, (size_t)L->host_size);
    for (int l = 0; l < 4; ++l) {
        int64_t s = L->dev_stride[l];
        uint8_t* base = bp + L->dev_base[l];
        for (int64_t g = 0; g < s; ++g) {
            uint32_t w[8];
            for (int j = 0; j < 8; ++j) {
                int64_t r = 8 * g + j;
                w[j] = r < rc[l] ? fwd(qc[L->n4 + rs[l] + r], mode) : 0u;
            }
            for (int k = 0; k < 32; ++k) {
                uint32_t byte = 0;
                for (int j = 0; j < 8; ++j) byte |= ((w[j] >> (31 - k)) & 1u) << (7 - j);
                base[k * s + g] = (uint8_t)byte;
            }
        }
    }
}

static void inverse_impl(const uint8_t* bp, const po_layout* L, int mode, const int32_t* begin,
                         const int32_t* end, int32_t* qc) {
    int64_t rs[4], rc[4];
    level_segments(L, rs, rc);
    for (int l = 0; l < 4; ++l) {
        int64_t s = L->dev_stride[l];
        const uint8_t* base = bp + L->dev_base[l];
        uint32_t planes = 0xFFFFFFFFu; /* bit k set = plane k used */
        if (begin) {
            int lr = 3 - l;
            planes = 0;
            if (begin[lr] < end[lr]) {
                for (int k = begin[lr]; k < end[lr] && k < 32; ++k) planes |= 1u << k;
                if (mode == 1) planes |= 1u; /* sign plane always, bitplane.cu:467-471 */
            }
        }
        for (int64_t r = 0; r < rc[l]; ++r) {
            int64_t g = r >> 3;
            int j = (int)(r & 7);
            uint32_t a = 0;
            for (int k = 0; k < 32; ++k)
                if (planes & (1u << k)) a |= (uint32_t)((base[k * s + g] >> (7 - j)) & 1u) << (31 - k);
            qc[L->n4 + rs[l] + r] = inv(a, mode);
        }
    }
}

void po_inverse_bitshuffle(const uint8_t* bp, const po_layout* L, int mode, int32_t* qc) {
    inverse_impl(bp, L, mode, NULL, NULL, qc);
}

void po_inverse_bitshuffle_window(const uint8_t* bp, const po_layout* L, int mode,
                                  const int32_t begin[4], const int32_t end[4], int32_t* qc) {
    inverse_impl(bp, L, mode, begin, end, qc);
}
