/*
 * strategy.c -- progressive load strategy of the CPU oracle (test
 * infrastructure only; see prism_oracle.h).  Restates
 * findOptimalStrategy / loadall / findStrategy_h, dataloader.cu:21-208.
 */
#include "prism_oracle.h"

#include <math.h>

/* Error cost of keeping only planes [0, b) of level lr, in units of eb:
 * ceil(2^(32-b) * F_lr) with F = sum of p^(12-3lr-{1,2,3}), p = 1 for lr 0
 * and 1.25 otherwise (dataloader.cu:59-69).  The reference evaluates this
 * with pow() in double; every F is an exact dyadic rational, so the exact
 * integer form below gives the same numbers. */
uint64_t po_strategy_cost(int lr, int b) {
    static const uint64_t num[4] = {3, 953125, 7625, 61};
    static const int den_log2[4] = {0, 16, 10, 4};
    if (b >= 32) return 0;
    int sh = 32 - b - den_log2[lr];
    if (sh >= 0) return num[lr] << sh;
    uint64_t d = (uint64_t)1 << (-sh);
    return (num[lr] + d - 1) / d;
}

static inline int rev_bits(int v, int nbits) {
    int r = 0;
    for (int i = 0; i < nbits; ++i) r |= ((v >> i) & 1) << (nbits - 1 - i);
    return r;
}

/* Exhaustive search over (b0..b3) in 0..32 maximising the skipped bytes
 * subject to the cost bound.  Ties resolve exactly as the reference's
 * reductions do: per thread (pos = tid + 1024k) the first strict maximum,
 * then the 1024-thread halving tree (keep left unless right is larger),
 * then the 64-slot block tree over b0.  A halving tree that keeps the left
 * operand on ties returns the tied index whose bit-reversed value is
 * smallest, so the winner is the maximum of
 *     (value, -rev6(b0), -rev10(tid), -k)
 * over all feasible combinations with value > 0; none -> {32,32,32,32}. */
void po_strategy_search(const uint64_t cum[128], uint64_t target_cost, int32_t choice[4],
                        uint64_t* best_value) {
    uint64_t value[4][33], cost[4][33];
    for (int lr = 0; lr < 4; ++lr) {
        value[lr][0] = cum[lr * 32 + 31]; /* cumulative quirk, dataloader.cu:47 */
        value[lr][32] = 0;
        for (int j = 31; j > 0; --j) value[lr][j] = cum[lr * 32 + 31] - cum[lr * 32 + j - 1];
        for (int b = 0; b <= 32; ++b) cost[lr][b] = po_strategy_cost(lr, b);
    }
    uint64_t bv = 0;
    int bt0 = 0, bt1 = 0, bt2 = 0;
    int have = 0;
    for (int b0 = 0; b0 <= 32; ++b0) {
        for (int pos = 0; pos < 33 * 33 * 33; ++pos) {
            int b1 = pos / (33 * 33), b2 = (pos / 33) % 33, b3 = pos % 33;
            uint64_t tc = cost[0][b0] + cost[1][b1] + cost[2][b2] + cost[3][b3];
            if (tc > target_cost) continue;
            uint64_t v = value[0][b0] + value[1][b1] + value[2][b2] + value[3][b3];
            if (v == 0) continue;
            int t0 = -rev_bits(b0, 6), t1 = -rev_bits(pos % 1024, 10), t2 = -(pos / 1024);
            int better = !have || v > bv || (v == bv && (t0 > bt0 || (t0 == bt0 && (t1 > bt1 || (t1 == bt1 && t2 > bt2)))));
            if (better) {
                have = 1;
                bv = v;
                bt0 = t0;
                bt1 = t1;
                bt2 = t2;
                choice[0] = b0;
                choice[1] = b1;
                choice[2] = b2;
                choice[3] = b3;
            }
        }
    }
    if (!have) {
        choice[0] = choice[1] = choice[2] = choice[3] = 32;
        bv = 0;
    }
    if (best_value) *best_value = bv;
}

int po_find_strategy(const uint64_t cum[128], int32_t begin[4], int32_t end[4], double eb,
                     double target_error, int* uses_search) {
    int maxlevel = 4; /* resetCounter, dataloader.cu:13-19 */
    if (target_error * 0.954 >= eb) {
        uint64_t target_cost = (uint64_t)floor(target_error * 0.954 / eb) - 1; /* 181 */
        int32_t ch[4];
        po_strategy_search(cum, target_cost, ch, NULL);
        for (int i = 0; i < 4; ++i) begin[i] = end[i];
        /* dataloader.cu:140-163: lr 3 first, so the last advancing lr wins */
        for (int lr = 3; lr >= 0; --lr) {
            if (begin[lr] < ch[lr]) {
                end[lr] = ch[lr];
                maxlevel = 3 - lr;
            } else {
                end[lr] = begin[lr];
            }
        }
        if (uses_search) *uses_search = 1;
    } else { /* loadall, dataloader.cu:21-28 */
        for (int i = 0; i < 4; ++i) begin[i] = end[i];
        for (int i = 0; i < 4; ++i) end[i] = 32;
        maxlevel = begin[0] < 32 ? 3 : begin[1] < 32 ? 2 : begin[2] < 32 ? 1 : begin[3] < 32 ? 0 : 4;
        if (uses_search) *uses_search = 0;
    }
    return maxlevel;
}
