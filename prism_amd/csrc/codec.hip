// codec.hip -- lossless bit-plane stream on gfx950: zero elimination plus
// three rounds of repetition elimination per 16 KiB chunk, chunk offsets by
// a single-pass decoupled look-back, the parallel decoder, and the trailer.
//
// Stream (zbpc.cu:155-263, RZE.hpp:10-167): u64 insize | u16 csize[C] |
// payloads in chunk order | zero pad to 8 | trailer.  Chunk order is level
// 3..0, plane 0..31, chunk index; chunk (l,b,c) covers host bytes
// [E_l + b*e_l + c*CS, min(+CS, E_l + (b+1)*e_l)).  Encoded chunk:
// [non-zero bytes][RE1 bytes][RE2 bytes][RE3 bytes][B3: 4 bytes][size: 2 B LE];
// bitmaps LSB-first; an all-zero chunk is the 2-byte footer alone; a chunk
// that does not shrink below its size is stored raw (csize == osize).
#include "common.hpp"

namespace pa {
namespace {

constexpr int kThreads = 256;
constexpr int kAvail = kCS - 2;

struct ChunkGeom {
    int64_t chunk_start[5];  // first chunk id of levels 3,2,1,0 (index 0..3), total at [4]
    int64_t cb[4];           // chunks per plane, by level
    int64_t e[4];            // host stride, by level
    int64_t hbase[4];        // host base, by level
    int64_t dbase[4];        // device base, by level
    int64_t dstride[4];      // device stride, by level
    int64_t nchunks, host_size;
};

ChunkGeom make_chunk_geom(const pa_layout* L) {
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

struct ChunkPos {
    int lv, plane;
    int64_t idx, base;
    int osize;
    bool last_of_plane;
};

__device__ __forceinline__ ChunkPos chunk_pos(const ChunkGeom& g, int64_t id) {
    int i = 0;
    while (i < 3 && id >= g.chunk_start[i + 1]) ++i;
    ChunkPos p;
    p.lv = 3 - i;
    const int64_t real = id - g.chunk_start[i];
    const int64_t cb = g.cb[p.lv];
    p.plane = (int)(real / cb);
    p.idx = real % cb;
    p.base = g.hbase[p.lv] + p.plane * g.e[p.lv] + p.idx * kCS;
    const int64_t rem = g.e[p.lv] - p.idx * kCS;
    p.osize = (int)(rem < kCS ? rem : kCS);
    p.last_of_plane = p.idx == cb - 1;
    return p;
}

// exclusive block scan of one int per lane (256 lanes = 4 wave64)
__device__ __forceinline__ int block_scan(int v, int* lds, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        const int s = lds[w];
        before += w < wave ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}

constexpr uint64_t kEmpty = 0, kAgg = 1, kPrefix = 2;

__device__ __forceinline__ void publish(uint64_t* status, int64_t id, uint64_t flag, uint64_t value) {
    __hip_atomic_store(status + id, (flag << 62) | value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back by wave 0: exclusive payload offset of chunk id.
// Chunk ids are handed out in arrival order, so every predecessor belongs to
// a workgroup that is already running and will publish without waiting.
__device__ int64_t look_back(uint64_t* status, int64_t id) {
    const int lane = threadIdx.x & 63;
    int64_t excl = 0, top = id - 1;
    while (true) {
        const int64_t idx = top - lane;
        const uint64_t st = idx >= 0 ? __hip_atomic_load(status + idx, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
                                     : (kPrefix << 62);
        const uint64_t flag = st >> 62;
        const uint64_t val = st & ((1ull << 62) - 1);
        const uint64_t pmask = __ballot(flag == kPrefix);
        const uint64_t emask = __ballot(flag == kEmpty);
        if (pmask) {
            const int p = __ffsll((unsigned long long)pmask) - 1;
            const uint64_t need = p == 63 ? ~0ull : ((2ull << p) - 1);
            if (emask & need) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            int64_t part = lane <= p ? (int64_t)val : 0;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
            return excl + part;
        }
        if (emask) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        int64_t part = (int64_t)val;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
        excl += part;
        top -= 64;
    }
}

// repetition elimination of 8 bytes (packed little endian) given the byte
// before them; returns the 8 flag bits
__device__ __forceinline__ uint32_t re_flags(uint64_t v, uint32_t prev) {
    uint32_t f = 0, p = prev;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t b = (uint32_t)(v >> (8 * j)) & 0xFFu;
        f |= (uint32_t)(b != p) << j;
        p = b;
    }
    return f;
}

__global__ void encode_init_kernel(uint64_t* status, int64_t n, uint32_t* counter) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        status[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0;
}

__global__ __launch_bounds__(kThreads) void encode_kernel(const uint8_t* __restrict__ bp, ChunkGeom g,
                                                          uint8_t* __restrict__ out, uint64_t* __restrict__ cum,
                                                          uint64_t* __restrict__ outsize,
                                                          uint64_t* __restrict__ status,
                                                          uint32_t* __restrict__ counter) {
    __shared__ uint64_t s_in[kCS / 8];
    __shared__ uint8_t s_out[kCS + 8];
    __shared__ uint64_t s_b0[256];
    __shared__ uint64_t s_b1[32];
    __shared__ uint8_t s_b2[32];
    __shared__ uint8_t s_b3[4];
    __shared__ int s_scan[8];
    __shared__ int64_t s_id, s_excl;
    const int tid = threadIdx.x;
    if (tid == 0) s_id = (int64_t)atomicAdd(counter, 1u);
    __syncthreads();
    const int64_t id = s_id;
    const ChunkPos cp = chunk_pos(g, id);
    const int osize = cp.osize;

    const uint64_t* src = reinterpret_cast<const uint64_t*>(bp + cp.base);
    const int nwords = osize >> 3;
    for (int i = tid; i < kCS / 8; i += kThreads) s_in[i] = i < nwords ? src[i] : 0ull;
    __syncthreads();

    // zero elimination: lane t owns bytes [64t, 64t+64)
    const uint8_t* inb = reinterpret_cast<const uint8_t*>(s_in);
    uint64_t mask = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        const uint64_t v = s_in[tid * 8 + w];
#pragma unroll
        for (int j = 0; j < 8; ++j) mask |= (uint64_t)(((v >> (8 * j)) & 0xFFu) != 0) << (8 * w + j);
    }
    int Z;
    int pos = block_scan(__popcll(mask), s_scan, &Z);
    s_b0[tid] = mask;
    bool fail = Z >= kAvail;
    if (!fail) {
        uint64_t m = mask;
        while (m) {
            const int j = __ffsll((unsigned long long)m) - 1;
            s_out[pos++] = inb[64 * tid + j];
            m &= m - 1;
        }
    }
    __syncthreads();
    int wpos = Z;
    int csize;
    if (Z == 0) {
        csize = 2;
    } else {
        // round 1: B0 (2048 bytes) -> B1 (256 bytes); lane t owns B0 bytes [8t, 8t+8)
        if (!fail) {
            const uint64_t cur = s_b0[tid];
            const uint32_t prev = tid ? (uint32_t)(s_b0[tid - 1] >> 56) : 0u;
            const uint32_t f = re_flags(cur, prev);
            reinterpret_cast<uint8_t*>(s_b1)[tid] = (uint8_t)f;
            int R;
            int p = block_scan(__popc(f), s_scan, &R);
            fail = wpos + R > kAvail;
            if (!fail) {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if ((f >> j) & 1u) s_out[wpos + p++] = (uint8_t)(cur >> (8 * j));
            }
            wpos += R;
            __syncthreads();
        }
        // round 2: B1 (256 bytes) -> B2 (32 bytes); lanes 0..31
        if (!fail) {
            uint64_t cur = 0;
            uint32_t f = 0;
            if (tid < 32) {
                cur = s_b1[tid];
                const uint32_t prev = tid ? (uint32_t)(s_b1[tid - 1] >> 56) : 0u;
                f = re_flags(cur, prev);
                s_b2[tid] = (uint8_t)f;
            }
            int R;
            int p = block_scan(__popc(f), s_scan, &R);
            fail = wpos + R > kAvail;
            if (!fail && tid < 32) {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if ((f >> j) & 1u) s_out[wpos + p++] = (uint8_t)(cur >> (8 * j));
            }
            wpos += R;
            __syncthreads();
        }
        // round 3: B2 (32 bytes) -> B3 (4 bytes); lanes 0..3
        if (!fail) {
            uint64_t cur = 0;
            uint32_t f = 0;
            if (tid < 4) {
                cur = reinterpret_cast<const uint64_t*>(s_b2)[tid];
                const uint32_t prev = tid ? (uint32_t)s_b2[8 * tid - 1] : 0u;
                f = re_flags(cur, prev);
                s_b3[tid] = (uint8_t)f;
            }
            int R;
            int p = block_scan(__popc(f), s_scan, &R);
            fail = wpos + R > kAvail;
            if (!fail && tid < 4) {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if ((f >> j) & 1u) s_out[wpos + p++] = (uint8_t)(cur >> (8 * j));
            }
            wpos += R;
            __syncthreads();
        }
        if (!fail && wpos >= kAvail - 4) fail = true;
        csize = fail ? osize : wpos + 6;
    }
    const bool raw = fail || csize >= osize;
    if (raw) csize = osize;
    if (!raw && tid == 0) {
        if (Z != 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) s_out[wpos + j] = s_b3[j];
        }
        s_out[csize - 2] = (uint8_t)(osize & 0xFF);
        s_out[csize - 1] = (uint8_t)((osize >> 8) & 0xFF);
    }

    // chunk offset
    if (tid == 0) publish(status, id, id == 0 ? kPrefix : kAgg, (uint64_t)csize);
    if (tid < 64) {
        const int64_t ex = id == 0 ? 0 : look_back(status, id);
        if (tid == 0) {
            s_excl = ex;
            if (id != 0) publish(status, id, kPrefix, (uint64_t)(ex + csize));
        }
    }
    __syncthreads();
    const int64_t excl = s_excl, incl = excl + csize;

    // payload store at an arbitrary byte offset: byte head, aligned words, byte tail
    const uint8_t* sb = raw ? inb : s_out;
    uint8_t* d = out + 8 + 2 * g.nchunks + excl;
    int head = (int)((4 - ((uintptr_t)d & 3)) & 3);
    if (head > csize) head = csize;
    if (tid < head) d[tid] = sb[tid];
    const int nw = (csize - head) >> 2;
    uint32_t* dw = reinterpret_cast<uint32_t*>(d + head);
    for (int w = tid; w < nw; w += kThreads) {
        const int k = head + 4 * w;
        dw[w] = (uint32_t)sb[k] | ((uint32_t)sb[k + 1] << 8) | ((uint32_t)sb[k + 2] << 16) |
                ((uint32_t)sb[k + 3] << 24);
    }
    const int tail0 = head + 4 * nw;
    if (tid < csize - tail0) d[tail0 + tid] = sb[tail0 + tid];

    if (tid == 0) {
        reinterpret_cast<uint16_t*>(out + 8)[id] = (uint16_t)csize;
        if (cp.last_of_plane) cum[(3 - cp.lv) * 32 + cp.plane] = (uint64_t)incl;
        if (id == g.nchunks - 1) {
            *reinterpret_cast<uint64_t*>(out) = (uint64_t)g.host_size;
            *outsize = (uint64_t)(8 + 2 * g.nchunks + incl);
        }
    }
}

// compress_merge: zero pad to 8 bytes, then flags, range, size table, anchors.
__global__ void merge_kernel(uint8_t* __restrict__ out, const uint64_t* __restrict__ outsize,
                             const uint8_t* __restrict__ flags8, const double* __restrict__ range,
                             const uint64_t* __restrict__ cum, const uint8_t* __restrict__ anchors,
                             int64_t anchor_bytes, uint64_t* __restrict__ total) {
    const uint64_t raw = *outsize;
    const uint64_t p = (raw + 7) & ~7ull;
    uint8_t* t = out + p;
    if (blockIdx.x == 0) {
        if (threadIdx.x < p - raw) out[raw + threadIdx.x] = 0;
        if (threadIdx.x < 8) t[threadIdx.x] = flags8[threadIdx.x];
        if (threadIdx.x == 0) {
            *reinterpret_cast<double*>(t + 8) = range ? *range : 1.0;
            *total = p + 1040 + (uint64_t)anchor_bytes;
        }
        for (int i = threadIdx.x; i < 128; i += blockDim.x) reinterpret_cast<uint64_t*>(t + 16)[i] = cum[i];
    }
    // anchors: 8-byte multiples (n4 * sizeof(T) with T of 4 bytes may leave 4)
    uint8_t* a = t + 1040;
    const int64_t nw = anchor_bytes >> 2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x)
        reinterpret_cast<uint32_t*>(a)[i] = reinterpret_cast<const uint32_t*>(anchors)[i];
}

// exclusive offsets of the payloads from the size table; one workgroup
__global__ __launch_bounds__(1024) void size_scan_kernel(const uint16_t* __restrict__ sizes, int64_t n,
                                                         int64_t* __restrict__ offsets) {
    __shared__ int64_t s_w[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t per = (n + 1023) / 1024;
    const int64_t lo = tid * per, hi = lo + per < n ? lo + per : n;
    int64_t sum = 0;
    for (int64_t i = lo; i < hi; ++i) sum += sizes[i];
    int64_t x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    int64_t before = 0;
    for (int w = 0; w < wave; ++w) before += s_w[w];
    int64_t run = before + x - sum;
    for (int64_t i = lo; i < hi; ++i) {
        offsets[i] = run;
        run += sizes[i];
    }
}

// Decoder: one workgroup per chunk (chunks are independent once the
// offsets are known).  WINDOW keeps only chunks holding a byte of a
// requested device bit-plane.
template <bool WINDOW>
__global__ __launch_bounds__(kThreads) void decode_kernel(const uint8_t* __restrict__ in, ChunkGeom g,
                                                          const int64_t* __restrict__ offsets,
                                                          uint8_t* __restrict__ bp,
                                                          const int32_t* __restrict__ begin,
                                                          const int32_t* __restrict__ end,
                                                          uint64_t* __restrict__ bytes_read,
                                                          int32_t* __restrict__ err, int64_t in_len) {
    __shared__ uint32_t s_c[kCS / 4 + 2];
    __shared__ uint64_t s_b0[256];
    __shared__ uint8_t s_b1[256];
    __shared__ uint8_t s_b2[32];
    __shared__ int s_scan[8];
    const int tid = threadIdx.x;
    const int64_t id = blockIdx.x;
    const ChunkPos cp = chunk_pos(g, id);
    const int osize = cp.osize;
    if (WINDOW) {
        bool sel = false;
#pragma unroll
        for (int lv = 0; lv < 4; ++lv) {
            const int lr = 3 - lv;
            const int b = begin[lr], e = end[lr];
            if (b < e) {
                const int64_t lo = g.dbase[lv] + (int64_t)b * g.dstride[lv];
                const int64_t hi = g.dbase[lv] + (int64_t)e * g.dstride[lv];
                if (cp.base < hi && cp.base + osize > lo) sel = true;
            }
        }
        if (!sel) return;
    }
    const uint16_t* sizes = reinterpret_cast<const uint16_t*>(in + 8);
    const int csize = sizes[id];
    const int64_t start = 8 + 2 * g.nchunks + offsets[id];
    uint64_t* dst0 = reinterpret_cast<uint64_t*>(bp + cp.base);
    if (*reinterpret_cast<const uint64_t*>(in) != (uint64_t)g.host_size || start + csize + 3 > in_len) {
        if (tid == 0) atomicOr(err, 8);
        for (int i = tid; i < (osize >> 3); i += kThreads) dst0[i] = 0ull;
        return;
    }
    if (WINDOW && bytes_read && tid == 0) atomicAdd((unsigned long long*)bytes_read, (unsigned long long)csize);
    const uint8_t* cbg = in + start;
    uint64_t* dst = reinterpret_cast<uint64_t*>(bp + cp.base);
    const int nwords = osize >> 3;
    if (csize > osize || csize < 2) {
        if (tid == 0) atomicOr(err, 1);
        for (int i = tid; i < nwords; i += kThreads) dst[i] = 0ull;
        return;
    }
    const int shift = (int)((uintptr_t)cbg & 3);
    const uint32_t* aw = reinterpret_cast<const uint32_t*>(cbg - shift);
    const int cw = (shift + csize + 3) >> 2;
    for (int i = tid; i < cw; i += kThreads) s_c[i] = aw[i];
    __syncthreads();
    const uint8_t* cb = reinterpret_cast<const uint8_t*>(s_c) + shift;
    if (csize == osize) {  // stored raw
        for (int i = tid; i < nwords; i += kThreads) {
            uint64_t v = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) v |= (uint64_t)cb[8 * i + j] << (8 * j);
            dst[i] = v;
        }
        return;
    }
    const int rpos = csize - 2;
    const int orig = (int)cb[rpos] | ((int)cb[rpos + 1] << 8);
    bool bad = orig != osize;
    if (rpos == 0) {
        if (bad && tid == 0) atomicOr(err, 2);
        for (int i = tid; i < nwords; i += kThreads) dst[i] = 0ull;
        return;
    }
    auto rd = [&](int k) -> uint32_t { return (k >= 0 && k < csize) ? (uint32_t)cb[k] : 0u; };
    bad = bad || rpos < 4;
    const uint32_t f3 = rd(rpos - 4) | (rd(rpos - 3) << 8) | (rd(rpos - 2) << 16) | (rd(rpos - 1) << 24);
    const int n3 = __popc(f3);
    const int d3 = rpos - 4 - n3;
    if (tid < 32) {
        const int c = __popc(f3 & (tid == 31 ? 0xFFFFFFFFu : ((2u << tid) - 1)));
        s_b2[tid] = c ? (uint8_t)rd(d3 + c - 1) : 0;
    }
    __syncthreads();
    const uint64_t* f2 = reinterpret_cast<const uint64_t*>(s_b2);
    const int n2 = __popcll(f2[0]) + __popcll(f2[1]) + __popcll(f2[2]) + __popcll(f2[3]);
    const int d2 = d3 - n2;
    {
        const int w = tid >> 6, bit = tid & 63;
        int c = __popcll(f2[w] & (bit == 63 ? ~0ull : ((2ull << bit) - 1)));
        for (int k = 0; k < w; ++k) c += __popcll(f2[k]);
        s_b1[tid] = c ? (uint8_t)rd(d2 + c - 1) : 0;
    }
    __syncthreads();
    const uint32_t f1 = s_b1[tid];
    int n1;
    const int e1 = block_scan(__popc(f1), s_scan, &n1);
    const int d1 = d2 - n1;
    {
        uint64_t v = 0;
        int c = e1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if ((f1 >> j) & 1u) {
                ++c;
                v |= (uint64_t)rd(d1 + c - 1) << (8 * j);
            } else if (c > 0) {
                v |= (uint64_t)rd(d1 + c - 1) << (8 * j);
            }
        }
        s_b0[tid] = v;
    }
    __syncthreads();
    const uint64_t f0 = s_b0[tid];
    int n0;
    const int e0 = block_scan(__popcll(f0), s_scan, &n0);
    bad = bad || d3 < 0 || d2 < 0 || d1 < 0 || n0 != d1;
    if (bad && tid == 0) atomicOr(err, 4);
    int c = e0;
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        uint64_t v = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if ((f0 >> (8 * w + j)) & 1ull) {
                v |= (uint64_t)rd(c) << (8 * j);
                ++c;
            }
        }
        const int wi = tid * 8 + w;
        if (wi < nwords) dst[wi] = bad ? 0ull : v;
    }
}

}  // namespace
}  // namespace pa

using namespace pa;

extern "C" size_t pa_encode_workspace(const pa_layout* L) {
    if (!L) return 0;
    return (size_t)L->nchunks * 8 + 256;
}

extern "C" int pa_lossless_encode(const uint8_t* d_bp, const pa_layout* L, uint8_t* d_out, uint64_t* d_cum,
                                  uint64_t* d_outsize, void* d_ws, size_t ws_bytes, void* stream) {
    if (!L || !d_bp || !d_out || !d_cum || !d_outsize) return PA_ERR_ARG;
    if (!d_ws || ws_bytes < pa_encode_workspace(L)) return PA_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    ChunkGeom g = make_chunk_geom(L);
    uint64_t* status = static_cast<uint64_t*>(d_ws);
    uint32_t* counter = reinterpret_cast<uint32_t*>(status + L->nchunks);
    int blocks = (int)((L->nchunks + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(encode_init_kernel, dim3(blocks), dim3(256), 0, st, status, L->nchunks, counter);
    PA_LAUNCH_CHECK();
    hipLaunchKernelGGL(encode_kernel, dim3((unsigned)L->nchunks), dim3(kThreads), 0, st, d_bp, g, d_out, d_cum,
                       d_outsize, status, counter);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_compress_merge(uint8_t* d_out, const pa_layout* L, int dtype, const uint64_t* d_outsize,
                                 const uint8_t* d_flags8, const double* d_range, const uint64_t* d_cum,
                                 const void* d_anchors, uint64_t* d_total, void* stream) {
    if (!L || !d_out || !d_outsize || !d_flags8 || !d_cum || !d_anchors || !d_total) return PA_ERR_ARG;
    if (dtype != PA_F32 && dtype != PA_F64) return PA_ERR_ARG;
    const int64_t ab = L->n4 * dtype;
    int blocks = (int)((ab / 4 + 255) / 256);
    if (blocks < 1) blocks = 1;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(merge_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), d_out, d_outsize, d_flags8,
                       d_range, d_cum, static_cast<const uint8_t*>(d_anchors), ab, d_total);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

// workspace: int32 error word (+pad) | int64 offsets[nchunks]
extern "C" size_t pa_decode_workspace(const pa_layout* L) {
    if (!L) return 0;
    return 64 + (size_t)L->nchunks * 8;
}

static int decode_common(const uint8_t* d_in, int64_t in_len, const pa_layout* L, const int32_t* d_begin, const int32_t* d_end,
                         uint8_t* d_bp, uint64_t* d_bytes_read, bool window, bool offsets_ready, void* d_ws,
                         size_t ws_bytes, hipStream_t st) {
    if (!L || !d_in || !d_bp) return PA_ERR_ARG;
    if (in_len < 8 + 2 * L->nchunks + L->trailer_fixed) return PA_ERR_STREAM;
    if (!d_ws || ws_bytes < pa_decode_workspace(L)) return PA_ERR_WORKSPACE;
    ChunkGeom g = make_chunk_geom(L);
    int32_t* err = static_cast<int32_t*>(d_ws);
    int64_t* offsets = reinterpret_cast<int64_t*>(static_cast<uint8_t*>(d_ws) + 64);
    if (!offsets_ready) {
        hipLaunchKernelGGL(encode_init_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<uint64_t*>(d_ws), 8,
                           reinterpret_cast<uint32_t*>(d_ws));
        PA_LAUNCH_CHECK();
        hipLaunchKernelGGL(size_scan_kernel, dim3(1), dim3(1024), 0, st,
                           reinterpret_cast<const uint16_t*>(d_in + 8), L->nchunks, offsets);
        PA_LAUNCH_CHECK();
    }
    if (window)
        hipLaunchKernelGGL(decode_kernel<true>, dim3((unsigned)L->nchunks), dim3(kThreads), 0, st, d_in, g, offsets,
                           d_bp, d_begin, d_end, d_bytes_read, err, in_len);
    else
        hipLaunchKernelGGL(decode_kernel<false>, dim3((unsigned)L->nchunks), dim3(kThreads), 0, st, d_in, g,
                           offsets, d_bp, nullptr, nullptr, nullptr, err, in_len);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_lossless_decode(const uint8_t* d_in, int64_t in_len, const pa_layout* L, uint8_t* d_bp,
                                  void* d_ws, size_t ws_bytes, void* stream) {
    return decode_common(d_in, in_len, L, nullptr, nullptr, d_bp, nullptr, false, false, d_ws, ws_bytes,
                         as_stream(stream));
}

extern "C" int pa_lossless_decode_window(const uint8_t* d_in, int64_t in_len, const pa_layout* L,
                                         const int32_t* d_begin,
                                         const int32_t* d_end, uint8_t* d_bp, uint64_t* d_bytes_read,
                                         int offsets_ready, void* d_ws, size_t ws_bytes, void* stream) {
    if (!d_begin || !d_end) return PA_ERR_ARG;
    return decode_common(d_in, in_len, L, d_begin, d_end, d_bp, d_bytes_read, true, offsets_ready != 0, d_ws, ws_bytes,
                         as_stream(stream));
}
