// quality.hip -- error statistics of a reconstruction against its original
// (statistic, include/utils/analyze.hpp:5-70) in one fused pass on gfx950.
//
// Figures: L-infinity error and the first index where it occurs, the sum of
// squared errors, min and max of the original, then MSE, PSNR = 20 log10
// (range) - 10 log10(MSE), NRMSE = sqrt(MSE) / range and Max_RE = L-inf /
// range.  A NaN in either array gives an infinite L-infinity error.
// Workgroups write partials to their own slots and one workgroup
// folds them in a fixed order, so every figure is deterministic.
#include "common.hpp"

namespace pa {
namespace {

constexpr int kThreads = 256;
constexpr int kBlocks = 1024;

struct Part {
    double max_err;
    int64_t idx;
    double sse, mn, mx;
};

// A NaN error (a NaN in either array) counts as an infinite one, so that it
// wins the maximum instead of failing every comparison.
__device__ __forceinline__ double nan_as_inf(double e) { return e == e ? e : __builtin_inf(); }

__device__ __forceinline__ void fold(Part& a, const Part& b) {
    if (b.max_err > a.max_err || (b.max_err == a.max_err && b.idx < a.idx)) {
        a.max_err = b.max_err;
        a.idx = b.idx;
    }
    a.sse += b.sse;
    a.mn = b.mn < a.mn ? b.mn : a.mn;
    a.mx = b.mx > a.mx ? b.mx : a.mx;
}

__device__ Part shfl_xor(const Part& p, int d) {
    Part q;
    q.max_err = __shfl_xor(p.max_err, d, 64);
    q.idx = __shfl_xor(p.idx, d, 64);
    q.sse = __shfl_xor(p.sse, d, 64);
    q.mn = __shfl_xor(p.mn, d, 64);
    q.mx = __shfl_xor(p.mx, d, 64);
    return q;
}

// Workgroup b takes the contiguous element range [b * per, (b + 1) * per);
// lane t of it the elements b * per + t + 256 k.  Within a lane the squared
// errors are added in index order; lanes and waves by a fixed tree.
template <typename T>
__global__ __launch_bounds__(kThreads) void quality_kernel(const T* __restrict__ ori, const T* __restrict__ rec,
                                                           int64_t n, int64_t per, Part* __restrict__ parts) {
    __shared__ Part s[kThreads / 64];
    Part p{-1.0, 0, 0.0, __builtin_inf(), -__builtin_inf()};
    const int64_t lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    constexpr int kBatch = 8;
    int64_t i = lo + threadIdx.x;
    for (; i + (kBatch - 1) * kThreads < hi; i += kBatch * kThreads) {
        T o[kBatch], r[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            o[k] = ori[i + k * kThreads];
            r[k] = rec[i + k * kThreads];
        }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const double ov = (double)o[k], e = nan_as_inf(__builtin_fabs((double)r[k] - ov));
            if (e > p.max_err) {
                p.max_err = e;
                p.idx = i + k * kThreads;
            }
            p.sse += e * e;
            p.mn = ov < p.mn ? ov : p.mn;
            p.mx = ov > p.mx ? ov : p.mx;
        }
    }
    for (; i < hi; i += kThreads) {
        const double ov = (double)ori[i], e = nan_as_inf(__builtin_fabs((double)rec[i] - ov));
        if (e > p.max_err) {
            p.max_err = e;
            p.idx = i;
        }
        p.sse += e * e;
        p.mn = ov < p.mn ? ov : p.mn;
        p.mx = ov > p.mx ? ov : p.mx;
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) fold(p, shfl_xor(p, d));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = p;
    __syncthreads();
    if (threadIdx.x == 0) {
        Part q = s[0];
        for (int w = 1; w < kThreads / 64; ++w) fold(q, s[w]);
        parts[blockIdx.x] = q;
    }
}

// out: [0] L-inf error, [1] its first index, [2] sum of squared errors,
// [3] min, [4] max of the original, [5] MSE, [6] PSNR (dB), [7] NRMSE,
// [8] Max_RE
__global__ __launch_bounds__(kThreads) void quality_finish_kernel(const Part* __restrict__ parts, int nparts, int64_t n,
                                                                  double* __restrict__ out) {
    __shared__ Part s[kThreads];
    Part p{-1.0, 0, 0.0, __builtin_inf(), -__builtin_inf()};
    for (int i = threadIdx.x; i < nparts; i += kThreads) fold(p, parts[i]);
    s[threadIdx.x] = p;
    __syncthreads();
    for (int h = kThreads / 2; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) fold(s[threadIdx.x], s[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const Part q = s[0];
        const double range = q.mx - q.mn, mse = q.sse / (double)n;
        out[0] = q.max_err;
        out[1] = (double)q.idx;
        out[2] = q.sse;
        out[3] = q.mn;
        out[4] = q.mx;
        out[5] = mse;
        out[6] = 20.0 * log10(range) - 10.0 * log10(mse);
        out[7] = sqrt(mse) / range;
        out[8] = q.max_err / range;
    }
}

}  // namespace
}  // namespace pa

using namespace pa;

extern "C" size_t pa_quality_workspace(int64_t n) {
    (void)n;
    return sizeof(Part) * kBlocks;
}

extern "C" int pa_quality(const void* d_ori, const void* d_rec, int dtype, int64_t n, double* d_out, void* d_ws,
                          size_t ws_bytes, void* stream) {
    if (!d_ori || !d_rec || !d_out || n <= 0 || (dtype != PA_F32 && dtype != PA_F64)) return PA_ERR_ARG;
    if (!d_ws || ws_bytes < pa_quality_workspace(n)) return PA_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    Part* parts = static_cast<Part*>(d_ws);
    const int64_t per = ((n + kBlocks - 1) / kBlocks + kThreads - 1) / kThreads * kThreads;
    const int blocks = (int)((n + per - 1) / per);
    if (dtype == PA_F32)
        hipLaunchKernelGGL(quality_kernel<float>, dim3(blocks), dim3(kThreads), 0, st, (const float*)d_ori,
                           (const float*)d_rec, n, per, parts);
    else
        hipLaunchKernelGGL(quality_kernel<double>, dim3(blocks), dim3(kThreads), 0, st, (const double*)d_ori,
                           (const double*)d_rec, n, per, parts);
    PA_LAUNCH_CHECK();
    hipLaunchKernelGGL(quality_finish_kernel, dim3(1), dim3(kThreads), 0, st, parts, blocks, n, d_out);
    PA_LAUNCH_CHECK();
    return PA_OK;
}
