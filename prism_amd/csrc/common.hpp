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

__host__ __device__ inline Geom make_geom(const pa_layout* L) {
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

// ------------------------------------------------------------------ batch --
// One field of a batch: its input, its layout, and the device buffers every
// stage of the compression reads and writes.  The batched kernels find the
// field of a workgroup in a per-stage prefix of workgroup counts.
struct BatchField {
    const void* data;
    int dtype;     // PA_F32 / PA_F64
    int mode;      // PA_MODE_NB / PA_MODE_SM
    int rel;       // 1: eb is relative to the value range
    double eb;     // relative or absolute bound
    double alpha;  // level scaling of the relative bound (alpha_of(rel eb); 1 in absolute mode)
    pa_layout L;
    Geom g;  // make_geom(&L)
    // auto-tuner sampling (host-computed; tune_ok = 0 leaves all flags false)
    int tune_ok, tsx0, tsy0, tsz0, tnsx, tnsy, tnblocks;
    // code-array ranges the predictor does not write (zeroed)
    int64_t gap_lo[5], gap_hi[5];
    // device buffers
    double* minmax;     // [min, max] + two 64-bit order keys
    double* ebinfo;     // eb_abs, range, eb_r[4] f64, ebx2[4] f64, then eb_r[4] | ebx2[4] f32
    uint8_t* flags;     // 8
    double* scores;     // 15
    double* tune_part;  // tnblocks x 15
    int32_t* qc;        // code_len
    void* anchors;      // n4 values
    uint8_t* bp;        // host_size
    int64_t* offsets;   // nchunks
    uint8_t* staging;   // nchunks x 16 KiB
    uint8_t* out;       // pa_compress_bound bytes
    uint64_t* cum;      // 128
    uint64_t* outsize;  // unpadded stream length
    uint64_t* result;   // [total stream bytes, status, range (f64 bits), absolute eb (f64 bits)]
};

// One progressive-retrieval round of one field in a batch of rounds: the
// field's stream and resident state, the round's target, and where the
// new reconstruction goes.
struct RetrJob {
    pa_layout L;
    Geom g;  // make_geom(&L)
    int dtype;
    int search;            // strategy: 1 = search under target_cost, 0 = load every plane
    uint64_t target_cost;
    int offsets_ready;     // decode offsets of an earlier round are in the workspace
    int first;             // first round: reconstruction from the anchors, else delta on old
    double le64[8];        // level bounds eb_r[4], ebx2[4] of a float64 field
    float le32[8];         // of a float32 field
    const uint8_t* stream;
    int64_t stream_len;
    const uint64_t* cum;   // trailer: cumulative payload sizes [4][32]
    const uint8_t* flags;  // trailer: flags8
    const void* anchors;   // trailer: anchors
    int32_t* state;        // begin[4], end[4], maxlevel
    unsigned long long* keys;  // strategy workgroup keys (33 x 33)
    uint8_t* bp;           // decoded bit-planes (host layout), kept between rounds
    int32_t* qc;           // codes, code_len
    uint8_t* dws;          // decode workspace: error word (64-byte slot) | offsets
    uint64_t* bytes_read;  // may be null
    const void* old;
    void* out;
};

// field of global work item b among fields [lo, hi): the largest f with
// prefix[f] <= b (fields with no work items share their successor's prefix)
__device__ __forceinline__ int find_field(const int64_t* prefix, int lo, int hi, int64_t b) {
    --hi;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
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
