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

// Compute units of the current device (cached per device; 256 on MI355X).
inline int device_cus() {
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cache[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev] = n;
    }
    return cache[dev];
}

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

// ------------------------------------------------------- stage geometry --
// Chunk grid of the codec: 16 KiB chunks of the host bit-plane layout, in
// id order levels 3..0, plane, index (zbpc.cu:155-263).
struct ChunkGeom {
    int64_t chunk_start[5];  // first chunk id of levels 3,2,1,0 (index 0..3), total at [4]
    int64_t cb[4];           // chunks per plane, by level
    int64_t e[4];            // host stride, by level
    int64_t hbase[4];        // host base, by level
    int64_t dbase[4];        // device base, by level
    int64_t dstride[4];      // device stride, by level
    int64_t nchunks, host_size;
};

__host__ __device__ inline ChunkGeom make_chunk_geom(const pa_layout* L) {
    ChunkGeom c;
    int64_t id = 0;
    for (int lv = 3, i = 0; lv >= 0; --lv, ++i) {
        c.chunk_start[i] = id;
        id += 32 * L->chunks_per_plane[lv];
    }
    c.chunk_start[4] = id;
    for (int lv = 0; lv < 4; ++lv) {
        c.cb[lv] = L->chunks_per_plane[lv];
        c.e[lv] = L->host_stride[lv];
        c.hbase[lv] = L->host_base[lv];
        c.dbase[lv] = L->dev_base[lv];
        c.dstride[lv] = L->dev_stride[lv];
    }
    c.nchunks = L->nchunks;
    c.host_size = L->host_size;
    return c;
}

// Level segments of the bit-plane shuffle (8-code groups; kShuffleGroups
// groups per workgroup).
constexpr int kShuffleGroups = 1024;
struct Levels {
    int64_t prefix[4];     // absolute start of each level's codes
    int64_t real_cnt[4];   // real codes per level
    int64_t seg_len[4];    // codes per level segment (pads and phantom tail included)
    int64_t stride[4];     // s_l
    int64_t base[4];       // A_l
    int blk_start[5];      // first workgroup of each level (levels 3,2,1,0 in order)
    int64_t dev_size, host_size;
};

__host__ __device__ inline Levels make_levels(const pa_layout* L) {
    Levels v;
    const int64_t cnt[4] = {L->n - L->cnt[0], L->cnt[0] - L->cnt[1], L->cnt[1] - L->cnt[2], L->cnt[2] - L->n4};
    int b = 0;
    for (int lv = 3, i = 0; lv >= 0; --lv, ++i) {
        v.prefix[lv] = L->prefix[lv];
        v.real_cnt[lv] = cnt[lv];
        v.seg_len[lv] = L->dev_len[lv];
        v.stride[lv] = L->dev_stride[lv];
        v.base[lv] = L->dev_base[lv];
        v.blk_start[i] = b;
        b += (int)((L->dev_stride[lv] + kShuffleGroups - 1) / kShuffleGroups);
    }
    v.blk_start[4] = b;
    v.dev_size = L->dev_size;
    v.host_size = L->host_size;
    return v;
}

// A progressive round whose window holds at most kPlaneWindow planes on
// every level is reconstructed straight from the decoded bit-planes: the
// deferring inverse shuffle then writes no codes.  Both kernels evaluate
// this on the same device window.
constexpr int kPlaneWindow = 8;
__device__ __forceinline__ bool window_from_planes(const int32_t* begin, const int32_t* end) {
    bool ok = true;
#pragma unroll
    for (int lr = 0; lr < 4; ++lr) ok = ok && end[lr] - begin[lr] <= kPlaneWindow;
    return ok;
}

// ------------------------------------------------------------------ batch --
// One field of a batch: its input, its layout, and the device buffers every
// stage of the compression reads and writes.  The batched kernels find the
// field of a workgroup in a per-stage prefix of workgroup counts.  The stage
// geometry is computed on the host: built in a kernel, its arrays would be
// indexed dynamically from private (scratch) memory.
struct BatchField {
    const void* data;
    int dtype;     // PA_F32 / PA_F64
    int mode;      // PA_MODE_NB / PA_MODE_SM
    int rel;       // 1: eb is relative to the value range
    double eb;     // relative or absolute bound
    double alpha;  // level scaling of the relative bound (alpha_of(rel eb); 1 in absolute mode)
    pa_layout L;
    Geom g;         // make_geom(&L)
    ChunkGeom cg;   // make_chunk_geom(&L)
    Levels lv;      // make_levels(&L)
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
    uint64_t* encws;    // encoder workspace: error word, then nchunks look-back status words
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
    Geom g;        // make_geom(&L)
    ChunkGeom cg;  // make_chunk_geom(&L)
    Levels lv;     // make_levels(&L)
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
    // begin[4], end[4], maxlevel, code source (0 int32, 1 int16, 2 bit-planes),
    // then at [10..13] each level's first live plane of the round's window
    int32_t* state;
    unsigned long long* keys;  // strategy workgroup keys (33 x 33)
    uint8_t* bp;           // decoded bit-planes (device layout in host_size bytes), kept between rounds
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
// The code array is shorter than 2^31 (layout creation rejects longer
// ones), so the index is evaluated in 32-bit unsigned arithmetic: every
// partial sum is below 2^32 and the result below 2^31.
template <int LEVEL>
__device__ __forceinline__ uint32_t lm_index(const Geom& g, int gx, int gy, int gz) {
    const uint32_t a = (uint32_t)gx >> LEVEL, b = (uint32_t)gy >> LEVEL, c = (uint32_t)gz >> LEVEL;
    uint32_t idx = a + b * (uint32_t)g.lx[LEVEL] + c * (uint32_t)g.lxy[LEVEL] + (uint32_t)g.prefix[LEVEL] -
                   ((c + 1) >> 1) * (uint32_t)g.lxy[LEVEL + 1];
    if ((c & 1) == 0) {
        idx -= ((b + 1) >> 1) * (uint32_t)g.lx[LEVEL + 1];
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
