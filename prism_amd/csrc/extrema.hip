// extrema.hip -- exact min / max of a field (StatBuffer::findMinMax,
// compressor.cu:228-302) on gfx950.
//
// One streaming pass with 16-byte loads; wave64 (DPP) then workgroup reduction,
// folded into two 64-bit words with atomicMin / atomicMax on an
// order-preserving integer image of the double value.  Min and max are
// exact, so the result does not depend on the reduction order.
#include "batch.hpp"
#include "common.hpp"
#include "extrema.hpp"

namespace pa {
namespace {

__global__ void extrema_init_kernel(unsigned long long* keys) {
    keys[0] = ~0ull;  // running min
    keys[1] = 0ull;   // running max
}

template <typename T>
__global__ __launch_bounds__(256) void extrema_kernel(const T* __restrict__ in, int64_t n,
                                                      unsigned long long* __restrict__ keys) {
    __shared__ double s_min[4], s_max[4];
    constexpr int V = 16 / sizeof(T);
    using Vec = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    double mn = __builtin_inf(), mx = -__builtin_inf();
    const int64_t nv = n / V;
    const Vec* vin = reinterpret_cast<const Vec*>(in);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // Each workgroup streams one contiguous chunk of whole 32 KiB batches;
    // a batch puts kBatch independent 16-byte loads per lane in flight
    // before any is reduced.
    constexpr int kBatch = 8;
    constexpr int64_t kSpan = (int64_t)kBatch * 256;  // vectors per batch
    const int64_t batches = nv / kSpan;
    const int64_t per = (batches + gridDim.x - 1) / gridDim.x;
    const int64_t b0 = blockIdx.x * per, b1 = b0 + per < batches ? b0 + per : batches;
    for (int64_t bt = b0; bt < b1; ++bt) {
        const Vec* p = vin + bt * kSpan + threadIdx.x;
        Vec v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) v[k] = p[k * 256];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const T* e = reinterpret_cast<const T*>(&v[k]);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                mn = fmin(mn, (double)e[j]);
                mx = fmax(mx, (double)e[j]);
            }
        }
    }
    for (int64_t i = batches * kSpan + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += stride) {
        const Vec v = vin[i];
        const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            mn = fmin(mn, (double)e[j]);
            mx = fmax(mx, (double)e[j]);
        }
    }
    for (int64_t i = nv * V + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        mn = fmin(mn, (double)in[i]);
        mx = fmax(mx, (double)in[i]);
    }
    wave_minmax(mn, mx);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_min[wave] = mn;
        s_max[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            mn = fmin(mn, s_min[w]);
            mx = fmax(mx, s_max[w]);
        }
        // atomics on one address from every workgroup serialise (about 12 ns
        // each on MI355X); a workgroup that cannot improve the running value
        // skips its atomic, and the running values settle after a few
        const unsigned long long kmn = order_key(mn), kmx = order_key(mx);
        if (kmn < __hip_atomic_load(keys, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(keys, kmn);
        if (kmx > __hip_atomic_load(keys + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(keys + 1, kmx);
    }
}

__global__ void extrema_finish_kernel(const unsigned long long* keys, double* out) {
    out[0] = key_value(keys[0]);
    out[1] = key_value(keys[1]);
}

}  // namespace
}  // namespace pa

using namespace pa;

extern "C" int pa_extrema(const void* d_data, int dtype, int64_t n, double* d_minmax, void* stream) {
    if (!d_data || !d_minmax || n <= 0 || (dtype != PA_F32 && dtype != PA_F64)) return PA_ERR_ARG;
    if (((uintptr_t)d_data & 15) != 0) return PA_ERR_ARG;
    hipStream_t st = as_stream(stream);
    // two 64-bit keys kept right after the two doubles of the result
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(d_minmax + 2);
    hipLaunchKernelGGL(extrema_init_kernel, dim3(1), dim3(1), 0, st, keys);
    PA_LAUNCH_CHECK();
    // 1024 workgroups (4 per CU) streaming contiguous ranges read faster on
    // MI355X than 2048 or 4096 (78 us against 90 us for 302 MB)
    const int64_t work = (n + 255) / 256;
    const int blocks = (int)(work < 1024 ? work : 1024);
    if (dtype == PA_F32)
        hipLaunchKernelGGL(extrema_kernel<float>, dim3(blocks), dim3(256), 0, st, (const float*)d_data, n, keys);
    else
        hipLaunchKernelGGL(extrema_kernel<double>, dim3(blocks), dim3(256), 0, st, (const double*)d_data, n, keys);
    PA_LAUNCH_CHECK();
    hipLaunchKernelGGL(extrema_finish_kernel, dim3(1), dim3(1), 0, st, keys, d_minmax);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int64_t pa::batch_ext_blocks(const pa_layout* L, int dtype, int ext_batches) {
    const int64_t nv = L->n / (16 / dtype);      // 16-byte vectors of the field
    return nv / (ext_batches * kExtBatch) + 1;  // the last workgroup also takes the scalar tail
}

