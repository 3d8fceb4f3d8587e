// strategy.hip -- progressive load strategy (which bit-planes of each level
// to read for a target error) on gfx950.
//
// The reference (dataloader.cu:30-177) enumerates the 33^4 plane counts with
// 33 blocks x 1024 threads, keeps each thread's first strict maximum, then
// reduces with halving trees that keep the left operand on ties.  Such a tree
// returns the tied index with the smallest bit-reversed value, so the same
// winner is the maximum of one 64-bit key
//     value << 22 | (63 - rev6(b0)) << 16 | (1023 - rev10(tid)) << 6 | (63 - k)
// (pos = tid + 1024k), which any reduction order finds.  A field's search is
// one launch of 33 workgroups (one per b0) whose best keys meet in one
// 64-bit slot by atomic maxima; the last to arrive updates the state.  The
// batched rounds write one key per (b0, b1) workgroup and reduce them in a
// second launch.
#include <rocprim/warp/warp_reduce.hpp>

#include "batch.hpp"
#include "common.hpp"

namespace pa {
namespace {

// ceil(2^(32-b) * F_lr), F = sum of p^(12-3lr-{1,2,3}), p = 1 (lr 0) or 1.25:
// the exact integers of the reference's double evaluation (dataloader.cu:59-68)
__device__ __forceinline__ uint64_t plane_cost(int lr, int b) {
    const uint64_t num = lr == 0 ? 3u : (lr == 1 ? 953125u : (lr == 2 ? 7625u : 61u));
    const int den = lr == 0 ? 0 : (lr == 1 ? 16 : (lr == 2 ? 10 : 4));
    if (b >= 32) return 0;
    const int sh = 32 - b - den;
    if (sh >= 0) return num << sh;
    const uint64_t d = 1ull << (-sh);
    return (num + d - 1) >> (-sh);  // the ceiling of num / d, d a power of two
}

// begin <- end; end <- max(end, choice); maxlevel = largest level that
// advanced (dataloader.cu:135-163), 4 when none did
__device__ void strategy_apply(unsigned long long key, int32_t* begin, int32_t* end, int32_t* maxlevel) {
    int ch[4] = {32, 32, 32, 32};
    if (key) {
        const int k = 63 - (int)(key & 63);
        const int t = (int)(__brev(1023u - (uint32_t)((key >> 6) & 1023)) >> 22);
        const int b0 = (int)(__brev(63u - (uint32_t)((key >> 16) & 63)) >> 26);
        const int pos = t + 1024 * k;
        ch[0] = b0;
        ch[1] = pos / 1089;
        ch[2] = (pos / 33) % 33;
        ch[3] = pos % 33;
    }
    int ml = 4;
    for (int i = 0; i < 4; ++i) begin[i] = end[i];
    for (int lr = 3; lr >= 0; --lr) {
        if (begin[lr] < ch[lr]) {
            end[lr] = ch[lr];
            ml = 3 - lr;
        } else {
            end[lr] = begin[lr];
        }
    }
    if (maxlevel) *maxlevel = ml;
}

constexpr int kSearchBlocks = 33 * 33;

// maximum over the 64 lanes of a wave (DPP row shifts, rocPRIM's cross-lane
// reduction); valid in lane 0
__device__ __forceinline__ unsigned long long wave_max(unsigned long long v) {
    using Reduce = rocprim::warp_reduce<unsigned long long, 64>;
    typename Reduce::storage_type storage;
    unsigned long long out;
    Reduce().reduce(v, out, storage, rocprim::maximum<unsigned long long>());
    return out;
}

// Grid (b0, b1) = 33 x 33 workgroups; lane p of a workgroup scores b2 =
// p / 33, b3 = p mod 33 (p < 1089) and the workgroup writes its best key to
// its own slot (no shared atomics: they serialise across workgroups).
__device__ __forceinline__ void strategy_search(const uint64_t* __restrict__ cum, uint64_t target_cost,
                                                unsigned long long* __restrict__ keys, int b0, int b1) {
    __shared__ uint64_t s_val[4][33], s_cost[4][33];
    __shared__ unsigned long long s_best[4];
    const int tid = threadIdx.x;
    if (tid < 4 * 33) {
        const int lr = tid / 33, b = tid % 33;
        uint64_t v;
        if (b == 0) v = cum[lr * 32 + 31];  // cumulative quirk, dataloader.cu:47
        else if (b == 32) v = 0;
        else v = cum[lr * 32 + 31] - cum[lr * 32 + b - 1];
        s_val[lr][b] = v;
        s_cost[lr][b] = plane_cost(lr, b);
    }
    __syncthreads();
    const uint64_t c01 = s_cost[0][b0] + s_cost[1][b1], v01 = s_val[0][b0] + s_val[1][b1];
    const uint64_t tie0 = (uint64_t)(63u - (__brev((uint32_t)b0) >> 26)) << 16;
    unsigned long long best_key = 0ull;
    if (c01 <= target_cost) {
        for (int p = tid; p < 33 * 33; p += 256) {
            const int b2 = p / 33, b3 = p % 33;
            const uint64_t tc = c01 + s_cost[2][b2] + s_cost[3][b3];
            if (tc > target_cost) continue;
            const uint64_t v = v01 + s_val[2][b2] + s_val[3][b3];
            if (v == 0) continue;
            const int pos = b1 * 1089 + p;  // the reference's thread + 1024 k position
            const uint64_t key = (v << 22) | tie0 |
                                 ((uint64_t)(1023u - (__brev((uint32_t)(pos & 1023)) >> 22)) << 6) |
                                 (uint64_t)(63 - (pos >> 10));
            if (key > best_key) best_key = key;
        }
    }
    best_key = wave_max(best_key);
    if ((tid & 63) == 0) s_best[tid >> 6] = best_key;
    __syncthreads();
    if (tid == 0) {
        unsigned long long k = s_best[0];
        for (int w = 1; w < 4; ++w) k = s_best[w] > k ? s_best[w] : k;
        keys[b1 * 33 + b0] = k;
    }
}

// One launch: workgroup b0 (of 33) scores, with lane l of 1024, the (b2, b3)
// = (p / 33, p mod 33) of p = l and l + 1024 (p < 1089) for every b1 (16
// waves per workgroup: the loop's latency is hidden), and takes its best
// key into *slot by an atomic maximum.  The arrival count
// (acquire-release) tells the 33rd workgroup that every maximum is in: it
// takes the slot, sets slot and count back to zero (zero before the first
// call, zero after every call) and updates begin / end / maxlevel.
constexpr int kRowThreads = 1024, kRowPer = (33 * 33 + kRowThreads - 1) / kRowThreads;
__global__ __launch_bounds__(kRowThreads) void strategy_kernel(const uint64_t* __restrict__ cum,
                                                               uint64_t target_cost, unsigned long long* slot,
                                                               uint32_t* count, int32_t* begin, int32_t* end,
                                                               int32_t* maxlevel) {
    __shared__ uint64_t s_val[4][33], s_cost[4][33];
    __shared__ unsigned long long s_best[kRowThreads / 64];
    const int tid = threadIdx.x, b0 = blockIdx.x;
    if (tid < 4 * 33) {
        const int lr = tid / 33, b = tid % 33;
        uint64_t v;
        if (b == 0) v = cum[lr * 32 + 31];  // cumulative quirk, dataloader.cu:47
        else if (b == 32) v = 0;
        else v = cum[lr * 32 + 31] - cum[lr * 32 + b - 1];
        s_val[lr][b] = v;
        s_cost[lr][b] = plane_cost(lr, b);
    }
    __syncthreads();
    uint64_t c23[kRowPer], v23[kRowPer];
#pragma unroll
    for (int i = 0; i < kRowPer; ++i) {
        const int p = tid + kRowThreads * i;
        c23[i] = 0;
        v23[i] = 0;
        if (p < 33 * 33) {
            c23[i] = s_cost[2][p / 33] + s_cost[3][p % 33];
            v23[i] = s_val[2][p / 33] + s_val[3][p % 33];
        }
    }
    const uint64_t tie0 = (uint64_t)(63u - (__brev((uint32_t)b0) >> 26)) << 16;
    unsigned long long best = 0ull;
    for (int b1 = 0; b1 < 33; ++b1) {
        const uint64_t c01 = s_cost[0][b0] + s_cost[1][b1];
        if (c01 > target_cost) continue;
        const uint64_t v01 = s_val[0][b0] + s_val[1][b1];
#pragma unroll
        for (int i = 0; i < kRowPer; ++i) {
            if (tid + kRowThreads * i >= 33 * 33 || c01 + c23[i] > target_cost) continue;
            const uint64_t v = v01 + v23[i];
            if (v == 0) continue;
            const int pos = b1 * 1089 + tid + kRowThreads * i;  // the reference's thread + 1024 k position
            const uint64_t key = (v << 22) | tie0 |
                                 ((uint64_t)(1023u - (__brev((uint32_t)(pos & 1023)) >> 22)) << 6) |
                                 (uint64_t)(63 - (pos >> 10));
            if (key > best) best = key;
        }
    }
    best = wave_max(best);
    if ((tid & 63) == 0) s_best[tid >> 6] = best;
    __syncthreads();
    if (tid == 0) {
        unsigned long long k = s_best[0];
        for (int w = 1; w < kRowThreads / 64; ++w) k = s_best[w] > k ? s_best[w] : k;
        __hip_atomic_fetch_max(slot, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 32u) {
            const unsigned long long kmax =
                __hip_atomic_exchange(slot, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            strategy_apply(kmax, begin, end, maxlevel);
        }
    }
}

// maximum of the workgroup keys, then the update of begin / end / maxlevel
__device__ __forceinline__ void strategy_reduce_apply(const unsigned long long* __restrict__ keys, int32_t* begin,
                                                      int32_t* end, int32_t* maxlevel) {
    __shared__ unsigned long long s_best[4];
    unsigned long long k = 0ull;
    for (int i = threadIdx.x; i < kSearchBlocks; i += 256) k = keys[i] > k ? keys[i] : k;
    k = wave_max(k);
    if ((threadIdx.x & 63) == 0) s_best[threadIdx.x >> 6] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) k = s_best[w] > k ? s_best[w] : k;
        strategy_apply(k, begin, end, maxlevel);
    }
}

// loadall (dataloader.cu:21-28)
__device__ __forceinline__ void load_all(int32_t* begin, int32_t* end, int32_t* maxlevel) {
    for (int i = 0; i < 4; ++i) begin[i] = end[i];
    for (int i = 0; i < 4; ++i) end[i] = 32;
    const int ml = begin[0] < 32 ? 3 : begin[1] < 32 ? 2 : begin[2] < 32 ? 1 : begin[3] < 32 ? 0 : 4;
    if (maxlevel) *maxlevel = ml;
}

__global__ void load_all_kernel(int32_t* begin, int32_t* end, int32_t* maxlevel) {
    if (threadIdx.x != 0) return;
    load_all(begin, end, maxlevel);
}

// batched rounds: grid (33, 33, jobs); jobs that load every plane skip the search
__global__ __launch_bounds__(256) void retr_search_kernel(const RetrJob* __restrict__ jobs) {
    const RetrJob& J = jobs[blockIdx.z];
    if (!J.search) return;
    strategy_search(J.cum, J.target_cost, J.keys, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(256) void retr_apply_kernel(const RetrJob* __restrict__ jobs) {
    const RetrJob& J = jobs[blockIdx.x];
    if (J.search) strategy_reduce_apply(J.keys, J.state, J.state + 4, J.state + 8);
    else if (threadIdx.x == 0) load_all(J.state, J.state + 4, J.state + 8);
}

}  // namespace
}  // namespace pa

using namespace pa;

// the best key's slot and the arrival count (zero before the first call)
extern "C" size_t pa_strategy_workspace(void) { return 2 * sizeof(unsigned long long); }

extern "C" int pa_find_strategy(const uint64_t* d_cum, int32_t* d_begin, int32_t* d_end, int32_t* d_maxlevel,
                                double eb, double target_error, void* d_ws, size_t ws_bytes, void* stream) {
    if (!d_cum || !d_begin || !d_end || !(eb > 0)) return PA_ERR_ARG;
    if (!d_ws || ws_bytes < pa_strategy_workspace()) return PA_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    if (target_error * 0.954 >= eb) {  // findStrategy_h 181, 193
        const double q = target_error * 0.954 / eb;
        if (!(q < 1.8e19)) return PA_ERR_ARG;
        const uint64_t target_cost = (uint64_t)__builtin_floor(q) - 1;
        unsigned long long* slot = static_cast<unsigned long long*>(d_ws);
        uint32_t* count = reinterpret_cast<uint32_t*>(slot + 1);
        hipLaunchKernelGGL(strategy_kernel, dim3(33), dim3(kRowThreads), 0, st, d_cum, target_cost, slot, count, d_begin,
                           d_end, d_maxlevel);
        PA_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(load_all_kernel, dim3(1), dim3(64), 0, st, d_begin, d_end, d_maxlevel);
        PA_LAUNCH_CHECK();
    }
    return PA_OK;
}

int pa::retr_strategy(const RetrDev& r, hipStream_t st) {
    hipLaunchKernelGGL(retr_search_kernel, dim3(33, 33, (unsigned)r.nj), dim3(256), 0, st, r.jobs);
    PA_LAUNCH_CHECK();
    hipLaunchKernelGGL(retr_apply_kernel, dim3((unsigned)r.nj), dim3(256), 0, st, r.jobs);
    PA_LAUNCH_CHECK();
    return PA_OK;
}
