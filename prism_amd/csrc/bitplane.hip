// bitplane.hip -- bit-plane shuffle and its inverses on gfx950.
//
// Layout (bitplane.cu:55-253, stated as a closed form): level l, plane k
// (0 = most significant bit) and 8-code group g of that level's segment map
// to byte A_l + k*s_l + g, whose bit 7-j is bit 31-k of code 8g+j (after the
// NB or SM transform).
//
// MI355X mapping: a lane owns 4 consecutive groups (32 codes) and turns them
// into one 32-bit word per plane with a single in-register 32x32 bit
// transpose, so each of the 32 plane rows is written with coalesced 4-byte
// stores (a wave covers 256 contiguous bytes of a row).  Codes are staged
// through LDS so the global reads are coalesced as well.  A workgroup of 256
// lanes covers 1024 groups (8192 codes) of one level.
#include "batch.hpp"
#include "common.hpp"

namespace pa {
namespace {

constexpr int kThreads = 256;
constexpr int kGroupsPerBlock = kShuffleGroups;       // 1024 groups
static_assert(kShuffleGroups == 4 * kThreads, "one group quartet per lane");
constexpr int kCodesPerBlock = 8 * kGroupsPerBlock;   // 8192 codes
constexpr int kRow = 33;                              // padded LDS row (32 codes)
constexpr int kStageWords = kThreads * kRow + 8;      // inverse: up to 3 codes of alignment shift

// In-register transpose of a 32x32 bit matrix: afterwards bit c of a[k]
// equals bit k of the original a[c] (bit 0 = least significant).
__device__ __forceinline__ void transpose32(uint32_t (&a)[32]) {
    uint32_t m = 0x0000FFFFu;
#pragma unroll
    for (int j = 16; j != 0; j >>= 1, m ^= (m << j)) {
#pragma unroll
        for (int k = 0; k < 32; k = ((k | j) + 1) & ~j) {
            const uint32_t t = ((a[k] >> j) ^ a[k + j]) & m;
            a[k] ^= t << j;
            a[k + j] ^= t;
        }
    }
}

// position of code q of a 4-group lane slice inside the transpose input:
// reversing codes within each group puts code j of a group at bit 7-j.
__device__ __forceinline__ int slot(int q) { return (q & ~7) | (7 - (q & 7)); }

__device__ __forceinline__ int block_level(const Levels& v, int blk, int* first) {
    int lv = 3, i = 0;
    while (i < 3 && blk >= v.blk_start[i + 1]) { ++i; --lv; }
    *first = v.blk_start[i];
    return lv;
}

template <int MODE>
__device__ __forceinline__ void shuffle_block(const int32_t* __restrict__ qc, const Levels& v, uint8_t* __restrict__ bp,
                                              int blk, int nblk, int32_t* stage) {
    int first;
    const int lv = block_level(v, blk, &first);
    const int64_t g0 = (int64_t)(blk - first) * kGroupsPerBlock;
    const int64_t c0 = 8 * g0;  // first code of the block within the level
    const int64_t real = v.real_cnt[lv];
    const int32_t* src = qc + v.prefix[lv];
    // all 32 loads of a lane are in flight before the first LDS write
    constexpr int kPer = kCodesPerBlock / kThreads;
    int32_t c[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t r = c0 + threadIdx.x + k * kThreads;
        c[k] = r < real ? src[r] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = threadIdx.x + k * kThreads;
        stage[(i >> 5) * kRow + (i & 31)] = c[k];
    }
    __syncthreads();
    const int64_t g = g0 + 4 * threadIdx.x;
    if (g < v.stride[lv]) {
        uint32_t a[32];
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const int32_t c = stage[threadIdx.x * kRow + q];
            a[slot(q)] = MODE == PA_MODE_NB ? nb_fwd(c) : sm_fwd(c);
        }
        transpose32(a);
        uint8_t* row = bp + v.base[lv] + g;
#pragma unroll
        for (int k = 0; k < 32; ++k) *reinterpret_cast<uint32_t*>(row + k * v.stride[lv]) = a[31 - k];
    }
    if (blk == nblk - 1) {  // bytes between the device and host layouts
        for (int64_t i = v.dev_size + 4 * threadIdx.x; i < v.host_size; i += 4 * kThreads)
            *reinterpret_cast<uint32_t*>(bp + i) = 0u;
    }
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void shuffle_kernel(const int32_t* __restrict__ qc, Levels v,
                                                           uint8_t* __restrict__ bp) {
    __shared__ int32_t stage[kThreads * kRow];
    shuffle_block<MODE>(qc, v, bp, blockIdx.x, gridDim.x, stage);
}

// one launch over the shuffle workgroups of every field of a batch
__global__ __launch_bounds__(kThreads) void batch_shuffle_kernel(const BatchField* __restrict__ fields,
                                                                 const int64_t* __restrict__ prefix, int nf) {
    __shared__ int32_t stage[kThreads * kRow];
    const int f = find_field(prefix, 0, nf, blockIdx.x);
    const BatchField& F = fields[f];
    const int blk = (int)(blockIdx.x - prefix[f]), nblk = (int)(prefix[f + 1] - prefix[f]);
    if (F.mode == PA_MODE_NB) shuffle_block<PA_MODE_NB>(F.qc, F.lv, F.bp, blk, nblk, stage);
    else shuffle_block<PA_MODE_SM>(F.qc, F.lv, F.bp, blk, nblk, stage);
}

// Inverse; WINDOW restricts each level to planes [begin[lr], end[lr]) (plus
// the sign plane in SM mode, bitplane.cu:467-471) and reads only those.
// Whether a window's codes fit int16: every level that reads planes starts
// at plane 17 or later, so its negabinary digits lie in bits 0..14 and the
// values in [-10922, 21845].
__device__ __forceinline__ bool window_narrow(const int32_t* begin, const int32_t* end) {
    bool narrow = true;
#pragma unroll
    for (int lr = 0; lr < 4; ++lr) narrow = narrow && (begin[lr] >= end[lr] || begin[lr] >= 17);
    return narrow;
}

// Stage slot of stage position d with one pad word per ROW words, so the
// lanes' writes of one instruction fall on distinct banks.
template <int ROW>
__device__ __forceinline__ int padded(int d) { return d + d / ROW; }

// Codes leave the stage as 16-byte stores: stage position d holds the code
// at destination index (block start - sh + d), sh the distance of the block
// start from the previous 16-byte boundary, so the vectors of 16 / E codes
// (E bytes each) are aligned; partial vectors at the block ends (and at the
// end of the level segment) are written code by code.
template <typename E, int ROW>
__device__ __forceinline__ void store_codes(const int32_t* stage, E* abase, int sh, int nvalid) {
    constexpr int V = 16 / (int)sizeof(E);  // codes per vector
    constexpr int W = V * (int)sizeof(E) / 4;  // 4 stage words per vector
    const int nv = (nvalid + sh + V - 1) / V;
    for (int vi = threadIdx.x; vi < nv; vi += kThreads) {
        const int lo = V * vi, hi = lo + V;
        if (lo >= sh && hi <= sh + nvalid) {
            const int p = padded<ROW>(W * vi);  // the 4 words share one row
            uint4 w;
            w.x = (uint32_t)stage[p];
            w.y = (uint32_t)stage[p + 1];
            w.z = (uint32_t)stage[p + 2];
            w.w = (uint32_t)stage[p + 3];
            *reinterpret_cast<uint4*>(abase + lo) = w;
        } else {
            const int a = lo > sh ? lo : sh, b = hi < sh + nvalid ? hi : sh + nvalid;
            for (int d = a; d < b; ++d) {
                if (sizeof(E) == 2)
                    abase[d] = (E)reinterpret_cast<const uint16_t*>(stage)[2 * padded<ROW>(d >> 1) + (d & 1)];
                else
                    abase[d] = (E)stage[padded<ROW>(d)];
            }
        }
    }
}

template <int MODE, bool WINDOW>
__device__ __forceinline__ void unshuffle_block(const uint8_t* __restrict__ bp, const Levels& v,
                                                const int32_t* __restrict__ begin, const int32_t* __restrict__ end,
                                                int32_t* __restrict__ qc, int blk, int32_t* stage,
                                                bool narrow = false) {
    int first;
    const int lv = block_level(v, blk, &first);
    const int64_t g0 = (int64_t)(blk - first) * kGroupsPerBlock;
    uint32_t planes = 0xFFFFFFFFu;
    if (WINDOW) {
        const int lr = 3 - lv;
        const int b = begin[lr], e = end[lr];
        planes = 0;
        if (b < e) {
            for (int k = b; k < e && k < 32; ++k) planes |= 1u << k;
            if (MODE == PA_MODE_SM) planes |= 1u;
        }
    }
    const int64_t g = g0 + 4 * threadIdx.x;
    const bool has = g < v.stride[lv];
    uint32_t a[32];
    if (has) {
        const uint8_t* row = bp + v.base[lv] + g;
#pragma unroll
        for (int k = 0; k < 32; ++k)  // planes are read once per round: no cache allocation
            a[31 - k] = (planes >> k) & 1u
                            ? __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(row + k * v.stride[lv]))
                            : 0u;
        transpose32(a);
    }
    const int64_t c0 = 8 * g0;
    const int64_t seg = v.seg_len[lv];
    const int nvalid = (int)(seg - c0 < kCodesPerBlock ? seg - c0 : kCodesPerBlock);  // <= 0: nothing to write
    if (narrow) {  // int16 codes, 16 words per stage row
        int16_t* base = reinterpret_cast<int16_t*>(qc) + v.prefix[lv] + c0;
        const int sh = (int)((reinterpret_cast<uintptr_t>(base) >> 1) & 7);
        if (has) {
            const int d0 = 32 * threadIdx.x + sh;
            if ((sh & 1) == 0) {  // code pairs share a stage word
#pragma unroll
                for (int q = 0; q < 32; q += 2) {
                    const uint32_t lo = (uint16_t)(MODE == PA_MODE_NB ? nb_inv(a[slot(q)]) : sm_inv(a[slot(q)]));
                    const uint32_t hi =
                        (uint16_t)(MODE == PA_MODE_NB ? nb_inv(a[slot(q + 1)]) : sm_inv(a[slot(q + 1)]));
                    stage[padded<16>((d0 + q) >> 1)] = (int32_t)(lo | (hi << 16));
                }
            } else {
                uint16_t* s16 = reinterpret_cast<uint16_t*>(stage);
#pragma unroll
                for (int q = 0; q < 32; ++q) {
                    const int d = d0 + q;
                    s16[2 * padded<16>(d >> 1) + (d & 1)] =
                        (uint16_t)(MODE == PA_MODE_NB ? nb_inv(a[slot(q)]) : sm_inv(a[slot(q)]));
                }
            }
        }
        __syncthreads();
        if (nvalid > 0) store_codes<int16_t, 16>(stage, base - sh, sh, nvalid);
        return;
    }
    int32_t* base = qc + v.prefix[lv] + c0;
    const int sh = (int)((reinterpret_cast<uintptr_t>(base) >> 2) & 3);
    if (has) {
        const int d0 = 32 * threadIdx.x + sh;
#pragma unroll
        for (int q = 0; q < 32; ++q) {
            const uint32_t w = a[slot(q)];
            stage[padded<32>(d0 + q)] = MODE == PA_MODE_NB ? nb_inv(w) : sm_inv(w);
        }
    }
    __syncthreads();
    if (nvalid > 0) store_codes<int32_t, 32>(stage, base - sh, sh, nvalid);
}

// NB window inverse that writes int16 codes when the window allows it and
// records the choice in *narrow for the reconstruction (0 int32, 1 int16).
// With `defer`, a window that the reconstruction reads from the bit-planes
// itself (window_from_planes) writes no codes and records 2.
__global__ __launch_bounds__(kThreads) void unshuffle_narrow_kernel(const uint8_t* __restrict__ bp, Levels v,
                                                                    const int32_t* __restrict__ begin,
                                                                    const int32_t* __restrict__ end,
                                                                    int32_t* __restrict__ qc,
                                                                    int32_t* __restrict__ narrow_out, int defer) {
    __shared__ int32_t stage[kStageWords];
    if (defer && window_from_planes(begin, end)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *narrow_out = 2;
        return;
    }
    const bool narrow = window_narrow(begin, end);
    if (blockIdx.x == 0 && threadIdx.x == 0) *narrow_out = narrow;
    // a deferring launch holds one resident grid, which takes the blocks in
    // turn (its workgroups then leave at once when the window is narrow)
    for (int blk = blockIdx.x; blk < v.blk_start[4]; blk += gridDim.x) {
        unshuffle_block<PA_MODE_NB, true>(bp, v, begin, end, qc, blk, stage, narrow);
        __syncthreads();  // the next block reuses the stage
    }
}

template <int MODE, bool WINDOW>
__global__ __launch_bounds__(kThreads) void unshuffle_kernel(const uint8_t* __restrict__ bp, Levels v,
                                                             const int32_t* __restrict__ begin,
                                                             const int32_t* __restrict__ end,
                                                             int32_t* __restrict__ qc) {
    __shared__ int32_t stage[kStageWords];
    unshuffle_block<MODE, WINDOW>(bp, v, begin, end, qc, blockIdx.x, stage);
}

// window inverse (NB) of every job of a batch of retrieval rounds
__global__ __launch_bounds__(kThreads) void retr_unshuffle_kernel(const RetrJob* __restrict__ jobs,
                                                                  const int64_t* __restrict__ prefix, int nj) {
    __shared__ int32_t stage[kStageWords];
    const int j = find_field(prefix, 0, nj, blockIdx.x);
    const RetrJob& J = jobs[j];
    const int blk = (int)(blockIdx.x - prefix[j]);
    const int32_t* live = J.state + 10;  // each level's first live plane (the planes before it are zero)
    if (J.g.nz > 1 && window_from_planes(live, J.state + 4)) {  // the reconstruction reads the planes
        if (blk == 0 && threadIdx.x == 0) J.state[9] = 2;
        return;
    }
    const bool narrow = window_narrow(live, J.state + 4);
    if (blk == 0 && threadIdx.x == 0) J.state[9] = narrow;
    unshuffle_block<PA_MODE_NB, true>(J.bp, J.lv, live, J.state + 4, J.qc, blk, stage, narrow);
}

}  // namespace
}  // namespace pa

using namespace pa;

extern "C" int pa_bitshuffle(const int32_t* d_qc, const pa_layout* L, int mode, uint8_t* d_bp, void* stream) {
    if (!L || !d_qc || !d_bp || (mode != PA_MODE_NB && mode != PA_MODE_SM)) return PA_ERR_ARG;
    Levels v = make_levels(L);
    hipStream_t st = as_stream(stream);
    if (mode == PA_MODE_NB)
        hipLaunchKernelGGL(shuffle_kernel<PA_MODE_NB>, dim3(v.blk_start[4]), dim3(kThreads), 0, st, d_qc, v, d_bp);
    else
        hipLaunchKernelGGL(shuffle_kernel<PA_MODE_SM>, dim3(v.blk_start[4]), dim3(kThreads), 0, st, d_qc, v, d_bp);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_inverse_bitshuffle(const uint8_t* d_bp, const pa_layout* L, int mode, int32_t* d_qc,
                                     void* stream) {
    if (!L || !d_qc || !d_bp || (mode != PA_MODE_NB && mode != PA_MODE_SM)) return PA_ERR_ARG;
    Levels v = make_levels(L);
    hipStream_t st = as_stream(stream);
    if (mode == PA_MODE_NB)
        hipLaunchKernelGGL((unshuffle_kernel<PA_MODE_NB, false>), dim3(v.blk_start[4]), dim3(kThreads), 0, st,
                           d_bp, v, nullptr, nullptr, d_qc);
    else
        hipLaunchKernelGGL((unshuffle_kernel<PA_MODE_SM, false>), dim3(v.blk_start[4]), dim3(kThreads), 0, st,
                           d_bp, v, nullptr, nullptr, d_qc);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_inverse_bitshuffle_window(const uint8_t* d_bp, const pa_layout* L, int mode,
                                            const int32_t* d_begin, const int32_t* d_end, int32_t* d_qc,
                                            void* stream) {
    if (!L || !d_qc || !d_bp || !d_begin || !d_end || (mode != PA_MODE_NB && mode != PA_MODE_SM))
        return PA_ERR_ARG;
    Levels v = make_levels(L);
    hipStream_t st = as_stream(stream);
    if (mode == PA_MODE_NB)
        hipLaunchKernelGGL((unshuffle_kernel<PA_MODE_NB, true>), dim3(v.blk_start[4]), dim3(kThreads), 0, st,
                           d_bp, v, d_begin, d_end, d_qc);
    else
        hipLaunchKernelGGL((unshuffle_kernel<PA_MODE_SM, true>), dim3(v.blk_start[4]), dim3(kThreads), 0, st,
                           d_bp, v, d_begin, d_end, d_qc);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

int64_t pa::batch_shuf_blocks(const pa_layout* L) { return make_levels(L).blk_start[4]; }

int pa::batch_shuffle(const BatchDev& b, hipStream_t st) {
    if (b.n_shuf <= 0) return PA_OK;
    hipLaunchKernelGGL(batch_shuffle_kernel, dim3((unsigned)b.n_shuf), dim3(kThreads), 0, st, b.fields, b.p_shuf, b.nf);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

int64_t pa::retr_unshuf_blocks(const pa_layout* L) { return make_levels(L).blk_start[4]; }

int pa::retr_unshuffle(const RetrDev& r, hipStream_t st) {
    if (r.n_unshuf <= 0) return PA_OK;
    hipLaunchKernelGGL(retr_unshuffle_kernel, dim3((unsigned)r.n_unshuf), dim3(kThreads), 0, st, r.jobs, r.p_unshuf,
                       r.nj);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_inverse_bitshuffle_window_narrow(const uint8_t* d_bp, const pa_layout* L, const int32_t* d_begin,
                                                   const int32_t* d_end, void* d_qc, int32_t* d_narrow,
                                                   void* stream) {
    if (!L || !d_qc || !d_bp || !d_begin || !d_end || !d_narrow) return PA_ERR_ARG;
    Levels v = make_levels(L);
    hipLaunchKernelGGL(unshuffle_narrow_kernel, dim3(v.blk_start[4]), dim3(kThreads), 0, as_stream(stream), d_bp, v,
                       d_begin, d_end, static_cast<int32_t*>(d_qc), d_narrow, 0);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_inverse_bitshuffle_window_deferred(const uint8_t* d_bp, const pa_layout* L,
                                                     const int32_t* d_begin, const int32_t* d_end, void* d_qc,
                                                     int32_t* d_src, void* stream) {
    if (!L || !d_qc || !d_bp || !d_begin || !d_end || !d_src) return PA_ERR_ARG;
    Levels v = make_levels(L);
    // only 3-D fields (nz > 1) reconstruct tile by tile from the planes; the
    // grid is what stays resident (four 33 KB workgroups per CU)
    const int defer = L->nz > 1 ? 1 : 0;
    const int resident = 4 * device_cus();
    const int grid = defer && v.blk_start[4] > resident ? resident : v.blk_start[4];
    hipLaunchKernelGGL(unshuffle_narrow_kernel, dim3(grid), dim3(kThreads), 0, as_stream(stream), d_bp, v, d_begin,
                       d_end, static_cast<int32_t*>(d_qc), d_src, defer);
    PA_LAUNCH_CHECK();
    return PA_OK;
}
