// common.hpp -- shared definitions of the HIP library: error handling, the
// field geometry passed to kernels, and small device helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prism_amd.h"

namespace pa {

constexpr int kCS = 16384;  // codec chunk size in bytes (include/encoder/lossless.hpp:4)
constexpr int kTS = 17;     // 16^3 tile plus one halo plane on the high side

#define PA_HIP_TRY(call)                        \
    do {                                        \
        hipError_t e_ = (call);                 \
        if (e_ != hipSuccess) return PA_ERR_HIP; \
    } while (0)

#define PA_LAUNCH_CHECK()                                   \
    do {                                                    \
        if (hipGetLastError() != hipSuccess) return PA_ERR_HIP; \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Geometry block handed to predictor kernels by value.
struct Geom {
    int nx, ny, nz;
    int gx, gy, gz;              // tiles per axis = anchor grid
    int64_t lx[5], lxy[5];       // level-grid leaps
    int64_t prefix[5];           // level-major region starts (absolute)
};

inline Geom make_geom(const pa_layout* L) {
    Geom g;
    g.nx = (int)L->nx;
    g.ny = (int)L->ny;
    g.nz = (int)L->nz;
    g.gx = (int)L->lx[4];
    g.gy = (int)(L->lxy[4] / L->lx[4]);
    g.gz = (int)(L->cnt[3] / L->lxy[4]);
    for (int i = 0; i < 5; ++i) {
        g.lx[i] = L->lx[i];
        g.lxy[i] = L->lxy[i];
        g.prefix[i] = L->prefix[i];
    }
    return g;
}

// Level-major position of a point of level LEVEL (< 4): its rank among the
// level's points in raster order, after the region start (compaction
// formula, InterpolationPredictor.cu:2070-2083).
template <int LEVEL>
__device__ __forceinline__ int64_t lm_index(const Geom& g, int gx, int gy, int gz) {
    const int64_t a = gx >> LEVEL, b = gy >> LEVEL, c = gz >> LEVEL;
    int64_t idx = a + b * g.lx[LEVEL] + c * g.lxy[LEVEL] + g.prefix[LEVEL] -
                  ((c + 1) >> 1) * g.lxy[LEVEL + 1];
    if ((c & 1) == 0) {
        idx -= ((b + 1) >> 1) * g.lx[LEVEL + 1];
        if ((b & 1) == 0) idx -= (a + 1) >> 1;
    }
    return idx;
}

__device__ __forceinline__ uint32_t nb_fwd(int32_t code) {
    return ((uint32_t)code + 0xAAAAAAAAu) ^ 0xAAAAAAAAu;
}
__device__ __forceinline__ int32_t nb_inv(uint32_t a) {
    return (int32_t)((a ^ 0xAAAAAAAAu) - 0xAAAAAAAAu);
}
__device__ __forceinline__ uint32_t sm_fwd(int32_t code) {
    return ((uint32_t)code & 0x7FFFFFFFu) ^ (uint32_t)(code >> 31);
}
__device__ __forceinline__ int32_t sm_inv(uint32_t a) {
    return (int32_t)((a & 0x7FFFFFFFu) ^ (uint32_t)(((int32_t)a) >> 31));
}

}  // namespace pa
