// extrema.hpp -- device pieces of the exact min / max (StatBuffer::findMinMax,
// compressor.cu:228-302) shared by the extrema kernels and the launch that
// runs the extrema beside the auto-tuner.
#pragma once
#include <rocprim/warp/warp_reduce.hpp>

#include "common.hpp"

namespace pa {

// order-preserving integer image of a double (and back): min / max become
// unsigned integer atomics
__device__ inline unsigned long long order_key(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double key_value(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)b);
}

constexpr int64_t kExtBatch = 8 * 256;  // vectors per load batch (8 per lane in flight)

struct FMin {
    __device__ double operator()(double a, double b) const { return fmin(a, b); }
};
struct FMax {
    __device__ double operator()(double a, double b) const { return fmax(a, b); }
};

// min / max over the 64 lanes of a wave with DPP row shifts (exact: the
// order of the comparisons does not matter); valid in every lane
__device__ inline void wave_minmax(double& mn, double& mx) {
    using Reduce = rocprim::warp_reduce<double, 64>;
    typename Reduce::storage_type storage;
    double a, b;
    Reduce().reduce(mn, a, storage, FMin());
    Reduce().reduce(mx, b, storage, FMax());
    mn = a;
    mx = b;
}

// Workgroup b of the batch streams `span` 16-byte vectors of its field in
// batches of 8 loads per lane (the field's tail elements go to its last
// workgroup); the running keys of each field live in its own descriptor.
template <typename T>
__device__ inline void extrema_span(const T* in, int64_t n, int64_t wg, int64_t span, unsigned long long* keys) {
    __shared__ double s_min[4], s_max[4];
    constexpr int V = 16 / sizeof(T);
    using Vec = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    const int64_t nv = n / V;
    const Vec* vin = reinterpret_cast<const Vec*>(in);
    double mn = __builtin_inf(), mx = -__builtin_inf();
    const int64_t v0 = wg * span;
    const int64_t v1 = v0 + span < nv ? v0 + span : nv;
    for (int64_t b = v0; b < v1; b += kExtBatch) {
        Vec v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t i = b + threadIdx.x + k * 256;
            if (i < v1) v[k] = vin[i];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t i = b + threadIdx.x + k * 256;
            if (i < v1) {
                const T* e = reinterpret_cast<const T*>(&v[k]);
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    mn = fmin(mn, (double)e[j]);
                    mx = fmax(mx, (double)e[j]);
                }
            }
        }
    }
    if (v0 + span >= nv) {  // last workgroup of the field: scalar tail
        for (int64_t i = nv * V + threadIdx.x; i < n; i += 256) {
            mn = fmin(mn, (double)in[i]);
            mx = fmax(mx, (double)in[i]);
        }
    }
    wave_minmax(mn, mx);
    if ((threadIdx.x & 63) == 0) {
        s_min[threadIdx.x >> 6] = mn;
        s_max[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            mn = fmin(mn, s_min[w]);
            mx = fmax(mx, s_max[w]);
        }
        const unsigned long long kmn = order_key(mn), kmx = order_key(mx);
        if (kmn < __hip_atomic_load(keys, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(keys, kmn);
        if (kmx > __hip_atomic_load(keys + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(keys + 1, kmx);
    }
}


}  // namespace pa
