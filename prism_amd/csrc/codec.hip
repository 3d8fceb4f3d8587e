// codec.hip -- lossless bit-plane stream on gfx950: zero elimination plus
// three rounds of repetition elimination per 16 KiB chunk (one chunk per
// wave64), each chunk's payload offset by a decoupled look-back over the
// chunks in dispatch order and its bytes written straight into the stream,
// the parallel decoder, and the trailer.  No pass takes an atomic
// read-modify-write on a shared address: on MI355X those serialise across
// workgroups at about 12 ns each.
//
// Stream (zbpc.cu:155-263, RZE.hpp:10-167): u64 insize | u16 csize[C] |
// payloads in chunk order | zero pad to 8 | trailer.  Chunk order is level
// 3..0, plane 0..31, chunk index; chunk (l,b,c) covers host bytes
// [E_l + b*e_l + c*CS, min(+CS, E_l + (b+1)*e_l)).  Encoded chunk:
// [non-zero bytes][RE1 bytes][RE2 bytes][RE3 bytes][B3: 4 bytes][size: 2 B LE];
// bitmaps LSB-first; an all-zero chunk is the 2-byte footer alone; a chunk
// that does not shrink below its size is stored raw (csize == osize).
#include <rocprim/warp/warp_scan.hpp>

#include <mutex>

#include "batch.hpp"
#include "common.hpp"

namespace pa {
namespace {

constexpr int kThreads = 256;
constexpr int kAvail = kCS - 2;

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
    // chunk ids and per-plane counts fit 32 bits (nchunks < 2^31)
    const uint32_t real = (uint32_t)(id - g.chunk_start[i]);
    const uint32_t cb = (uint32_t)g.cb[p.lv];
    p.plane = (int)(real / cb);
    p.idx = real - (uint32_t)p.plane * cb;
    p.base = g.hbase[p.lv] + p.plane * g.e[p.lv] + p.idx * kCS;
    const int64_t rem = g.e[p.lv] - p.idx * kCS;
    p.osize = (int)(rem < kCS ? rem : kCS);
    p.last_of_plane = p.idx == cb - 1;
    return p;
}

// inclusive scan over the 64 lanes of a wave with DPP row shifts and row
// broadcasts (rocPRIM's cross-lane form: no LDS round trip per step)
__device__ __forceinline__ int wave_incl_scan(int v) {
    using Scan = rocprim::warp_scan<int, 64>;
    typename Scan::storage_type storage;
    int out;
    Scan().inclusive_scan(v, out, storage);
    return out;
}

// exclusive block scan of one int per lane (256 lanes = 4 wave64)
__device__ __forceinline__ int block_scan(int v, int* lds, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = wave_incl_scan(v);
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


// 4-bit mask of the non-zero bytes of w
__device__ __forceinline__ uint32_t nz_bytes(uint32_t w) {
    const uint32_t h = (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
    return ((h >> 7) & 1u) | ((h >> 14) & 2u) | ((h >> 21) & 4u) | ((h >> 28) & 8u);
}

// keep the ordering of LDS accesses of one wave across phases
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); __builtin_amdgcn_wave_barrier(); }

// LDS of one wave: the zero-elimination bitmap, one step's bytes and the
// repetition-elimination bytes with the footer.  The surviving bytes go to
// the stream step by step, so a wave needs 5.5 KiB of LDS (an encoded chunk
// held whole would need 16 KiB and cap the encoder at two waves per SIMD).
constexpr int kTailBytes = 2048 + 256 + 32 + 4 + 2;  // RE1 + RE2 + RE3 kept bytes, B3, size
struct __align__(16) WaveSlot {
    uint16_t b0[kCS / 16];  // zero-elimination bitmap: 2 bytes per 16-byte segment
    // a step's bytes: the zero-eliminated form keeps the (< 4) bytes carried
    // from the previous step's last partial word in front of its surviving
    // bytes; a raw chunk puts the previous step's last word at 12..16 and
    // the step's 1 KiB from byte 16 (plus one word of slack)
    uint8_t step[16 + 1024 + 16];
    // the RE bytes, B3 and footer from byte 4, after up to 4 bytes that
    // carry the last partial word of the zero-eliminated bytes (plus slack
    // for the funnel reads of the last word)
    uint8_t tail[(4 + kTailBytes + 8 + 15) / 16 * 16];
    int zs[16];  // surviving bytes per step
};
static_assert(sizeof(WaveSlot) % 16 == 0, "slots must keep 16-byte alignment");

// Writes the 4-byte-aligned global words dst[0..nw) from the LDS byte
// sequence at src + d0 + 4i (any byte offset: funnel-shifted from two
// aligned LDS words), but only the bytes whose offset from dst lies in
// [lo, hi): a word shared with a neighbouring chunk is written byte by
// byte, so the two chunks' waves never overwrite each other's bytes.
__device__ __forceinline__ void store_funnel(const uint8_t* src, int d0, uint32_t* dst, int nw, int lo, int hi,
                                             int lane) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    for (int i = lane; i < nw; i += 64) {
        const int d = d0 + 4 * i;
        const uint32_t v = __builtin_amdgcn_alignbyte(s[(d >> 2) + 1], s[d >> 2], d & 3);
        const int x = 4 * i;
        if (x >= lo && x + 4 <= hi) {
            dst[i] = v;
        } else {
            uint8_t* a = reinterpret_cast<uint8_t*>(dst + i);
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (x + b >= lo && x + b < hi) a[b] = (uint8_t)(v >> (8 * b));
        }
    }
}

// ---- decoupled look-back over the encoder's workgroups in index order ----
// A workgroup encodes kW consecutive chunks (one per wave) and publishes one
// status word: flag in bits 62..63 (0: not yet published, 1: the compressed
// size of its chunks, 2: the inclusive payload offset after them) and the
// value in bits 0..61.  Workgroups are dispatched in index order, so every
// workgroup one waits for has been dispatched before it: the lowest
// unfinished workgroup never waits.  The wait is bounded all the same: a
// wave that polls kSpinLimit times records an error, takes the missing
// sizes as 0 (its offset can then only be too small, never past the
// stream's end) and goes on.
constexpr uint64_t kStAgg = 1ull << 62, kStInc = 2ull << 62, kStVal = (1ull << 62) - 1;
constexpr int kSpinLimit = 1 << 20;

__device__ __forceinline__ void status_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t status_load(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publishes the size `agg` of workgroup w's chunks and returns its exclusive
// payload offset (zbpc.cu:113-153), run by one wave.  A poll reads a window
// of 64 x kLbPer status words (lane l: distances kLbPer*l .. kLbPer*l +
// kLbPer - 1 below `top`); the sizes up to the nearest published offset are
// summed, and a window with no published offset adds all its sizes and the
// next poll reads the window below it.  Status words bypass the caches, so
// a poll that finds a size still missing backs off before it reads again:
// the polls of thousands of waiting waves would otherwise crowd out the
// chunk loads.
constexpr int kLbPer = 1;

__device__ __forceinline__ int64_t look_back(uint64_t* status, int64_t w, int64_t agg, int32_t* err) {
    const int lane = threadIdx.x & 63;
    if (w == 0) {
        if (lane == 0) status_store(status, kStInc | (uint64_t)agg);
        return 0;
    }
    if (lane == 0) status_store(status + w, kStAgg | (uint64_t)agg);
    int64_t excl = 0;
    int64_t top = w - 1;
    int polls = 0, backoff = 1;
    for (;;) {
        uint64_t s[kLbPer];
#pragma unroll
        for (int t = 0; t < kLbPer; ++t) {
            const int64_t j = top - kLbPer * lane - t;
            s[t] = j >= 0 ? status_load(status + j) : kStInc;  // before workgroup 0: offset 0
        }
        // this lane's entries nearest first: sizes before its first offset
        int first_t = kLbPer, lsum = 0;
        bool lmiss = false;
#pragma unroll
        for (int t = kLbPer - 1; t >= 0; --t) {
            const uint32_t fl = (uint32_t)(s[t] >> 62);
            if (fl == 2) {
                first_t = t;
                lsum = 0;
                lmiss = false;
            } else {
                lsum += fl == 1 ? (int)(s[t] & kStVal) : 0;
                lmiss = lmiss || fl == 0;
            }
        }
        const uint64_t inc = __ballot(first_t < kLbPer);
        const int first = inc ? (int)__builtin_ctzll(inc) : 64;  // lanes 0..first make up the sum
        const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1);
        const bool missing = (__ballot(lmiss) & need) != 0;
        if (missing && ++polls < kSpinLimit) {
            for (int b = 0; b < backoff; ++b) __builtin_amdgcn_s_sleep(4);
            backoff = backoff < 16 ? 2 * backoff : 16;
            continue;
        }
        if (missing && lane == 0) atomicOr(err, 1);
        // sizes (each below 2^18, at most 64 kLbPer of them: an int sum),
        // then the offset that lane `first` holds
        const int sum = __builtin_amdgcn_readlane(wave_incl_scan(lane <= first ? lsum : 0), 63);
        excl += sum;
        if (first < 64) {
            uint64_t v = 0;
#pragma unroll
            for (int t = 0; t < kLbPer; ++t) v = t == first_t ? (s[t] & kStVal) : v;
            const uint32_t vlo = __builtin_amdgcn_readlane((uint32_t)v, first);
            const uint32_t vhi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), first);
            excl += (int64_t)(((uint64_t)vhi << 32) | vlo);
            break;
        }
        top -= 64 * kLbPer;
    }
    if (lane == 0) status_store(status + w, kStInc | (uint64_t)(excl + agg));
    return excl;
}

// Repetition-elimination flags (RZE.hpp:10-86).  Round 1: lane l owns
// bytes [32l, 32l+32) of the zero-elimination bitmap B0; bit i of the
// result is set where B0 byte 32l+i differs from the byte before it (byte
// -1 = 0).
__device__ __forceinline__ uint32_t re1_flags(const uint16_t* b0, int lane) {
    const uint4* b0v = reinterpret_cast<const uint4*>(b0) + 2 * lane;
    const uint4 q0 = b0v[0], q1 = b0v[1];
    const uint32_t v[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    uint32_t prev = __shfl_up(v[7] >> 24, 1, 64);
    if (lane == 0) prev = 0;
    uint32_t f1 = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t shifted = (v[j] << 8) | (j == 0 ? prev : (v[j - 1] >> 24));
        f1 |= nz_bytes(v[j] ^ shifted) << (4 * j);
    }
    return f1;
}
// round 2: B1 = the round-1 flags as bytes (lane l holds bytes [4l, 4l+4))
__device__ __forceinline__ uint32_t re2_flags(uint32_t f1, int lane) {
    uint32_t p2 = __shfl_up(f1 >> 24, 1, 64);
    if (lane == 0) p2 = 0;
    return nz_bytes(f1 ^ ((f1 << 8) | p2));
}
// round 3: B2 byte m = the round-2 nibbles of lanes 2m, 2m+1 (m < 32)
__device__ __forceinline__ uint32_t re2_byte(uint32_t f2, int lane) {
    return (__shfl(f2, (2 * lane) & 63, 64) | (__shfl(f2, (2 * lane + 1) & 63, 64) << 4)) & 0xFFu;
}
__device__ __forceinline__ bool re3_keep(uint32_t f2, int lane) {
    const uint32_t b2 = re2_byte(f2, lane);
    uint32_t p3 = __shfl_up(b2, 1, 64);
    if (lane == 0) p3 = 0;
    return lane < 32 && b2 != p3;
}

// Where an encoder writes: the stream, its cumulative size table and
// length, the look-back status words and the error word.
struct EncodeOut {
    uint8_t* stream;
    uint64_t* cum;
    uint64_t* outsize;
    uint64_t* status;
    int32_t* err;
};

// a workgroup's chunk sizes and payload offset
struct GroupLds {
    int csize[8];
    int64_t off;
};

// One chunk per wave64, kW chunks per workgroup (chunks id0 .. id0 + kW - 1,
// workgroup w of the stream), in one pass: zero elimination (16 steps of
// 1 KiB; lane l owns 16-byte segment 64k+l, so every load is coalesced),
// the three repetition-elimination rounds on registers and shuffles, the
// compressed sizes, the workgroup's look-back for its payload offset, then
// the bytes straight to their place in the stream (zbpc.cu:155-263).  Every
// wave of the workgroup calls it (a wave past the last chunk has size 0).
template <int kW>
__device__ __forceinline__ void encode_chunk(const uint8_t* __restrict__ bp, const ChunkGeom& g, const EncodeOut& o,
                                             int64_t id0, int64_t w, WaveSlot& W, GroupLds& G) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t id = id0 + wave;
    const bool valid = id < g.nchunks;
    const ChunkPos cp = chunk_pos(g, valid ? id : 0);
    const int osize = valid ? cp.osize : 0;
    const uint8_t* src = bp + cp.base;

    // the whole chunk is requested before the first step (16 KiB per wave,
    // 64 registers per lane) so it costs one memory latency; it stays in
    // registers until its bytes are written
    constexpr int kSteps = kCS / 1024;
    uint2 lo8[kSteps], hi8[kSteps];
#pragma unroll
    for (int k = 0; k < kSteps; ++k) {
        const int b = 16 * (64 * k + lane);
        lo8[k] = make_uint2(0u, 0u);
        hi8[k] = make_uint2(0u, 0u);
        if (b < osize) {
            lo8[k] = *reinterpret_cast<const uint2*>(src + b);
            if (b + 8 < osize) hi8[k] = *reinterpret_cast<const uint2*>(src + b + 8);
        }
    }
    // zero-elimination bitmaps and the surviving-byte count
    int Z = 0;
#pragma unroll
    for (int k = 0; k < kSteps; ++k) {
        const uint32_t m = nz_bytes(lo8[k].x) | (nz_bytes(lo8[k].y) << 4) | (nz_bytes(hi8[k].x) << 8) |
                           (nz_bytes(hi8[k].y) << 12);
        W.b0[64 * k + lane] = (uint16_t)m;
        const int c = __builtin_amdgcn_readlane(wave_incl_scan(__popc(m)), 63);  // the wave's total
        if (lane == 0) W.zs[k] = c;
        Z += c;
    }
    bool fail = Z >= kAvail;
    wave_sync();
    // repetition elimination, counts only: the flags of rounds 1 and 2 stay
    // in two registers, the kept bytes are placed after the zero-eliminated
    // bytes have left (the chunk's registers are free by then)
    int wpos = Z;  // chunk position
    uint32_t f1 = 0, f2 = 0, b3 = 0;
    if (Z != 0 && !fail) {  // (a chunk that already failed is raw whatever follows)
        // round 1: B0 (2048 bytes) -> B1 (256 bytes); lane l owns B0 bytes [32l, 32l+32)
        f1 = re1_flags(W.b0, lane);
        wpos += __builtin_amdgcn_readlane(wave_incl_scan(__popc(f1)), 63);
        fail = fail || wpos > kAvail;
        // round 2: B1 (256 bytes, lane l holds bytes [4l, 4l+4) = f1) -> B2 (32 bytes)
        f2 = re2_flags(f1, lane);
        wpos += __builtin_amdgcn_readlane(wave_incl_scan(__popc(f2)), 63);
        fail = fail || wpos > kAvail;
        // round 3: B2 byte m = nibbles of lanes 2m, 2m+1 -> B3 (4 bytes)
        b3 = (uint32_t)__ballot(re3_keep(f2, lane));
        wpos += __popc(b3);
        fail = fail || wpos > kAvail || wpos >= kAvail - 4;
    }
    int csize = Z == 0 ? 2 : (fail ? osize : wpos + 6);
    const bool raw = fail || csize >= osize;
    if (raw) csize = osize;
    if (!valid) csize = 0;

    // the workgroup's payload offset, then the chunk's, the size table, the
    // per-plane cumulative sizes, the stream header and length
    if (lane == 0) G.csize[wave] = csize;
    __syncthreads();
    if (wave == 0) {
        int64_t agg = 0;
#pragma unroll
        for (int k = 0; k < kW; ++k) agg += G.csize[k];
        const int64_t goff = look_back(o.status, w, agg, o.err);
        if (lane == 0) G.off = goff;
    }
    __syncthreads();
    if (!valid) return;
    int64_t off = G.off;
#pragma unroll
    for (int k = 0; k < kW; ++k) off += k < wave ? G.csize[k] : 0;
    if (lane == 0) {
        reinterpret_cast<uint16_t*>(o.stream + 8)[id] = (uint16_t)csize;
        if (cp.last_of_plane) o.cum[(3 - cp.lv) * 32 + cp.plane] = (uint64_t)(off + csize);
        if (id == g.nchunks - 1) *o.outsize = (uint64_t)(8 + 2 * g.nchunks + off + csize);
        if (id == 0) *reinterpret_cast<uint64_t*>(o.stream) = (uint64_t)g.host_size;
    }
    uint8_t* P = o.stream + 8 + 2 * g.nchunks + off;  // first byte of the payload

    if (raw) {
        // step k: the previous step's last word at step[12..16), the step's
        // 1 KiB from step[16]; global word i of the step starts at chunk
        // position 1024k + 4i - q (q = misalignment of P), LDS byte
        // 16 + 4i - q
        const int q = (int)((uintptr_t)P & 3);
        uint32_t* dst = reinterpret_cast<uint32_t*>(P - q);
#pragma unroll
        for (int k = 0; k < kSteps; ++k) {
            if (1024 * k >= osize) break;  // wave-uniform
            const uint32_t prev_last = k > 0 ? __shfl(hi8[k - 1].y, 63, 64) : 0u;
            reinterpret_cast<uint4*>(W.step + 16)[lane] = make_uint4(lo8[k].x, lo8[k].y, hi8[k].x, hi8[k].y);
            if (lane == 0) reinterpret_cast<uint32_t*>(W.step)[3] = prev_last;
            wave_sync();
            const int rest = osize - 1024 * k;  // chunk bytes from this step on
            // words whose last byte lies in this step (the final word of the
            // chunk, when it ends past a word boundary, goes with the last step)
            const int nw = rest > 1024 ? 256 : (rest + q + 3) >> 2;
            store_funnel(W.step, 16 - q, dst + 256 * k, nw, q - 1024 * k, q + osize - 1024 * k, lane);
            wave_sync();  // the next step reuses the buffer
        }
        return;
    }
    // zero-eliminated bytes, step by step from the registers the bitmap
    // pass filled (the loop is unrolled so they stay in registers): a step's
    // bytes follow the r < 4 bytes carried from the previous step's last
    // partial word (r = alignment of the step's first byte in the stream),
    // whole words go out, the last partial word is carried on
    int zpos = 0;
    if (Z != 0) {
#pragma unroll
        for (int k = 0; k < kSteps; ++k) {
            const uint32_t w[4] = {lo8[k].x, lo8[k].y, hi8[k].x, hi8[k].y};
            const int zs = W.zs[k];
            if (zs == 0) continue;  // wave-uniform
            const int r = (int)((uintptr_t)(P + zpos) & 3);
            const uint32_t m = W.b0[64 * k + lane];
            const int cnt = __popc(m);
            int pos = r + wave_incl_scan(cnt) - cnt;
#pragma unroll
            for (int i = 0; i < 16; ++i)
                if ((m >> i) & 1u) W.step[pos++] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
            wave_sync();
            const int nfull = (r + zs) >> 2;
            store_funnel(W.step, 0, reinterpret_cast<uint32_t*>(P + zpos - r), nfull, r - zpos, r + csize - zpos, lane);
            wave_sync();  // the stores have read the buffer
            if (lane == 0) reinterpret_cast<uint32_t*>(W.step)[0] = reinterpret_cast<const uint32_t*>(W.step)[nfull];
            wave_sync();  // the next step reuses the buffer
            zpos += zs;
        }
    }
    // the RE bytes at tail[4 + (chunk position - Z)], B3 and the footer,
    // after the carried bytes (stream positions Z - rt .. Z - 1); then every
    // word from the aligned P + Z - rt to the end
    const int rt = (int)((uintptr_t)(P + Z) & 3);
    if (lane < rt) W.tail[4 - rt + lane] = W.step[lane];
    int pos = 4;
    if (Z != 0) {
        {  // round 1: B0 byte 32 lane + i is kept where bit i of f1 is set
            const int cnt = __popc(f1);
            int p = pos + wave_incl_scan(cnt) - cnt;
            const uint4* b0v = reinterpret_cast<const uint4*>(W.b0) + 2 * lane;
            const uint4 q0 = b0v[0], q1 = b0v[1];
            const uint32_t v[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
            for (int i = 0; i < 32; ++i)
                if ((f1 >> i) & 1u) W.tail[p++] = (uint8_t)(v[i >> 2] >> (8 * (i & 3)));
            pos += __builtin_amdgcn_readlane(wave_incl_scan(cnt), 63);
        }
        {  // round 2: B1 byte 4 lane + i (byte i of f1) where bit i of f2 is set
            const int cnt = __popc(f2);
            int p = pos + wave_incl_scan(cnt) - cnt;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if ((f2 >> i) & 1u) W.tail[p++] = (uint8_t)(f1 >> (8 * i));
            pos += __builtin_amdgcn_readlane(wave_incl_scan(cnt), 63);
        }
        {  // round 3: the B2 bytes where B3 is set
            const uint32_t b2 = re2_byte(f2, lane);
            if (lane < 32 && ((b3 >> lane) & 1u)) W.tail[pos + __popc(b3 & ((1u << lane) - 1u))] = (uint8_t)b2;
            pos += __popc(b3);
        }
        if (lane < 4) W.tail[pos + lane] = (uint8_t)(b3 >> (8 * lane));
        pos += 4;
    }
    if (lane == 0) {
        W.tail[pos] = (uint8_t)(osize & 0xFF);
        W.tail[pos + 1] = (uint8_t)((osize >> 8) & 0xFF);
    }
    wave_sync();
    store_funnel(W.tail, 4 - rt, reinterpret_cast<uint32_t*>(P + Z - rt), (rt + csize - Z + 3) >> 2, rt - Z,
                 rt + csize - Z, lane);
}

// one chunk per wave, kW chunks per workgroup
template <int kW>
__global__ __launch_bounds__(64 * kW) void encode_kernel(const uint8_t* __restrict__ bp, ChunkGeom g, EncodeOut o) {
    __shared__ WaveSlot slots[kW];
    __shared__ GroupLds G;
    encode_chunk<kW>(bp, g, o, (int64_t)blockIdx.x * kW, blockIdx.x, slots[threadIdx.x >> 6], G);
}

// one launch over the chunks of every field of a batch (one chunk per
// wave); each field's chunks start at a multiple of 8 in the launch (its
// prefix entry), so a workgroup never spans two fields, and they are
// dispatched in id order
template <int kW>
__global__ __launch_bounds__(64 * kW) void batch_encode_kernel(const BatchField* __restrict__ fields,
                                                               const int64_t* __restrict__ prefix, int nf) {
    __shared__ WaveSlot slots[kW];
    __shared__ GroupLds G;
    const int64_t gid0 = (int64_t)blockIdx.x * kW;
    const int f = find_field(prefix, 0, nf, gid0);
    const BatchField& F = fields[f];
    const EncodeOut o{F.out, F.cum, F.outsize, F.encws + 1, reinterpret_cast<int32_t*>(F.encws)};
    const int64_t id0 = gid0 - prefix[f];
    encode_chunk<kW>(F.bp, F.cg, o, id0, id0 / kW, slots[threadIdx.x >> 6], G);
}

// clears an encoder workspace: the error word, then the n status words
__global__ void encode_clear_kernel(uint64_t* __restrict__ ws, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) ws[i] = 0ull;
}

// the same for every field of a batch (one thread per launch slot of a
// chunk, plus one per field for its error word)
__global__ void batch_encode_clear_kernel(const BatchField* __restrict__ fields, const int64_t* __restrict__ prefix,
                                          int nf, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total + nf) return;
    const int64_t gid = i < total ? i : 0;
    const int f = i < total ? find_field(prefix, 0, nf, gid) : (int)(i - total);
    const BatchField& F = fields[f];
    const int64_t k = i < total ? gid - prefix[f] : -1;  // status word k (-1: the error word)
    if (k < F.cg.nchunks) F.encws[1 + k] = 0ull;
}

// The decoder's payload offsets by one workgroup scan of the size table
// (zbpc.cu:265-365 sums the sizes before each chunk).  One workgroup of 1024
// lanes takes the table in windows of kScanLds sizes: a window is staged
// through LDS with coalesced 8-byte loads, lane t sums its run of
// consecutive sizes, a wave then workgroup scan gives the run offsets, and
// the window's total carries into the next.
constexpr int kScanLds = 16384;

struct ScanLds {
    int64_t w[16];
    uint16_t sz[kScanLds];
};

__device__ __forceinline__ void scan_sizes(const uint16_t* __restrict__ sizes, int64_t n, ScanLds& S,
                                           int64_t* __restrict__ offsets) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool words = ((uintptr_t)sizes & 7) == 0;
    int64_t carry = 0;
    for (int64_t w0 = 0; w0 < n; w0 += kScanLds) {
        const int64_t m = n - w0 < kScanLds ? n - w0 : kScanLds;
        {
            // 8-byte words (4 sizes), all of a lane's loads in flight at once:
            // the table may have just been written by other CUs, so each
            // round trip is long
            const int64_t nw = words ? m / 4 : 0;
            const uint2* wv = reinterpret_cast<const uint2*>(sizes + w0);
            constexpr int kU = kScanLds / 4 / 1024;
            uint2 v[kU];
#pragma unroll
            for (int k = 0; k < kU; ++k) {
                const int64_t i = tid + 1024 * k;
                v[k] = i < nw ? wv[i] : make_uint2(0u, 0u);
            }
#pragma unroll
            for (int k = 0; k < kU; ++k) {
                const int64_t i = tid + 1024 * k;
                if (i < nw) reinterpret_cast<uint2*>(S.sz)[i] = v[k];
            }
            for (int64_t i = 4 * nw + tid; i < m; i += 1024) S.sz[i] = sizes[w0 + i];
        }
        __syncthreads();
        const int64_t per = (m + 1023) / 1024;
        const int64_t lo = tid * per, hi = lo + per < m ? lo + per : m;
        int64_t sum = 0;
        for (int64_t i = lo; i < hi; ++i) sum += S.sz[i];
        int64_t x = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) S.w[wave] = x;
        __syncthreads();
        int64_t run = carry + x - sum, tot = 0;
        for (int w = 0; w < 16; ++w) {
            run += w < wave ? S.w[w] : 0;
            tot += S.w[w];
        }
        for (int64_t i = lo; i < hi; ++i) {
            offsets[w0 + i] = run;
            run += S.sz[i];
        }
        carry += tot;
        __syncthreads();  // the next window reuses S
    }
}

// compress_merge body: zero pad to 8 bytes, then flags, range, size table,
// anchors; workgroup blk of nblk.
__device__ __forceinline__ void merge_block(uint8_t* __restrict__ out, const uint64_t* __restrict__ outsize,
                                            const uint8_t* __restrict__ flags8, const double* __restrict__ range,
                                            const uint64_t* __restrict__ cum, const uint8_t* __restrict__ anchors,
                                            int64_t anchor_bytes, uint64_t* __restrict__ total, int64_t blk,
                                            int64_t nblk) {
    const uint64_t raw = *outsize;
    const uint64_t p = (raw + 7) & ~7ull;
    uint8_t* t = out + p;
    if (blk == 0) {
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
    for (int64_t i = blk * blockDim.x + threadIdx.x; i < nw; i += nblk * blockDim.x)
        reinterpret_cast<uint32_t*>(a)[i] = reinterpret_cast<const uint32_t*>(anchors)[i];
}

// compress_merge: zero pad to 8 bytes, then flags, range, size table, anchors.
__global__ void merge_kernel(uint8_t* __restrict__ out, const uint64_t* __restrict__ outsize,
                             const uint8_t* __restrict__ flags8, const double* __restrict__ range,
                             const uint64_t* __restrict__ cum, const uint8_t* __restrict__ anchors,
                             int64_t anchor_bytes, uint64_t* __restrict__ total) {
    merge_block(out, outsize, flags8, range, cum, anchors, anchor_bytes, total, blockIdx.x, gridDim.x);
}

// trailer of every field of a batch; the range is the one the stream
// records (1.0 in absolute mode)
__global__ void batch_merge_kernel(const BatchField* __restrict__ fields, const int64_t* __restrict__ prefix, int nf) {
    const int f = find_field(prefix, 0, nf, blockIdx.x);
    const BatchField& F = fields[f];
    merge_block(F.out, F.outsize, F.flags, F.ebinfo + 1, F.cum, static_cast<const uint8_t*>(F.anchors),
                F.L.n4 * F.dtype, F.result, blockIdx.x - prefix[f], prefix[f + 1] - prefix[f]);
}

// exclusive offsets of the payloads from the size table; one workgroup
__device__ __forceinline__ void size_scan(const uint16_t* __restrict__ sizes, int64_t n, int64_t* __restrict__ offsets) {
    __shared__ ScanLds S;
    scan_sizes(sizes, n, S, offsets);
}

// also clears the decoder's error words (the 64-byte slot before the offsets)
__global__ __launch_bounds__(1024) void size_scan_kernel(const uint16_t* __restrict__ sizes, int64_t n,
                                                         int64_t* __restrict__ offsets, uint64_t* __restrict__ err) {
    if (threadIdx.x < 8) err[threadIdx.x] = 0ull;
    size_scan(sizes, n, offsets);
}

// Chunks tile the bit-plane buffer in id order (levels 3..0, plane, index),
// so the chunks overlapping one device window form one id range and a
// round's selection is at most four ranges in increasing order.
__device__ __forceinline__ int64_t chunk_of_byte(const ChunkGeom& g, int64_t x) {
    int i = 0;  // levels 3, 2, 1, 0 in increasing host offset
    while (i < 3 && x >= g.hbase[2 - i]) ++i;
    const int lv = 3 - i;
    const int64_t r = x - g.hbase[lv];
    // the host plane p = r / e < 32 by bisection (no 64-bit divide)
    int p = 0;
#pragma unroll
    for (int s = 16; s >= 1; s >>= 1)
        if (r >= (int64_t)(p + s) * g.e[lv]) p += s;
    return g.chunk_start[i] + p * g.cb[lv] + (r - p * g.e[lv]) / kCS;
}

struct Window {
    int64_t lo[4], hi[4];  // disjoint, increasing; [lo, hi) may be empty
    int64_t count;
};

// The chunks holding a byte of a requested device plane: for each level with
// begin < end, those overlapping [A_l + begin s_l, A_l + end s_l) (device
// strides are at least 8 bytes, so a window is never empty).  A chunk shared
// by two windows belongs to the first.
__device__ __forceinline__ Window window_ranges(const ChunkGeom& g, const int32_t* begin, const int32_t* end) {
    Window w;
    int64_t prev = 0;
    w.count = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {  // lr = i, level 3 - i: device windows in increasing offset
        const int lv = 3 - i;
        const int b = begin[i], e = end[i];
        int64_t lo = prev, hi = prev;
        if (b < e) {
            lo = chunk_of_byte(g, g.dbase[lv] + (int64_t)b * g.dstride[lv]);
            hi = chunk_of_byte(g, g.dbase[lv] + (int64_t)e * g.dstride[lv] - 1) + 1;
            lo = lo < prev ? prev : lo;
            hi = hi < lo ? lo : hi;
            prev = hi;
        }
        w.lo[i] = lo;
        w.hi[i] = hi;
        w.count += hi - lo;
    }
    return w;
}

// k-th selected chunk (k < count)
__device__ __forceinline__ int64_t window_chunk(const Window& w, int64_t k) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t n = w.hi[i] - w.lo[i];
        if (k < n) return w.lo[i] + k;
        k -= n;
    }
    return w.lo[3] + k;
}

// The first plane of each level's window [begin, end) that a covering chunk
// holds a non-zero byte of (an empty chunk is its 2-byte footer alone and
// decodes to zeros), end where there is none: the planes before it are zero
// and need neither decoding nor reading.  A plane is tested through the
// chunks that cover its device bytes, from their sizes and offsets.  Run by
// one wave (lane l tests planes l and 64 + l of the 4 x 32); valid in every
// lane.
__device__ __forceinline__ void live_begins(const ChunkGeom& g, const int64_t* offsets, const uint16_t* sizes,
                                            const int32_t* begin, const int32_t* end, int (&live)[4]) {
    const int lane = threadIdx.x & 63;
    uint64_t m[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = lane + 64 * h, lr = p >> 5, k = p & 31, lv = 3 - lr;
        bool on = false;
        if (k >= begin[lr] && k < end[lr]) {
            const int64_t s = g.dbase[lv] + (int64_t)k * g.dstride[lv];
            const int64_t lo = chunk_of_byte(g, s), hi = chunk_of_byte(g, s + g.dstride[lv] - 1);
            on = offsets[hi] + sizes[hi] - offsets[lo] != 2 * (hi - lo + 1);
        }
        m[h] = __ballot(on);
    }
#pragma unroll
    for (int lr = 0; lr < 4; ++lr) {
        const uint32_t bits = (uint32_t)(m[lr >> 1] >> (32 * (lr & 1)));
        live[lr] = bits ? __builtin_ctz(bits) : end[lr];
    }
}

// payload bytes of the selected chunks, from the exclusive offsets
__device__ __forceinline__ uint64_t window_bytes(const Window& w, const int64_t* offsets, const uint16_t* sizes) {
    uint64_t t = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (w.hi[i] > w.lo[i]) t += (uint64_t)(offsets[w.hi[i] - 1] + sizes[w.hi[i] - 1] - offsets[w.lo[i]]);
    return t;
}

// Decoder: one workgroup per chunk (chunks are independent once the
// offsets are known).  The windowed form walks the selected id ranges.
struct DecodeLds {
    uint32_t c[kCS / 4 + 4];  // payload; the last words are read past its end and discarded
    uint64_t b0[256];
    uint64_t sel[256];  // byte-permute selectors of each 8-bit zero mask
    int e0[256];        // first payload byte of each lane's 64 output bytes
    uint8_t b1[256];
    uint8_t b2[32];
    int scan[8];
};

// Selector of mask m for v_perm_b32 over 8 payload bytes: output byte j is
// payload byte popc(m & (2^j - 1)) when bit j of m is set, else 0 (selector
// 0x0C).  Written once per workgroup; read after the chunk's barriers.
__device__ __forceinline__ void init_selectors(DecodeLds& S) {
    const uint32_t m = threadIdx.x;
    uint64_t sel = 0;
    int c = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) sel |= (uint64_t)((m >> j) & 1u ? c++ : 0x0C) << (8 * j);
    S.sel[m] = sel;
}

// 8 payload bytes from byte position b of the LDS words
__device__ __forceinline__ void bytes8(const uint32_t* w, int b, uint32_t& lo, uint32_t& hi) {
    const int q = b >> 2, o = b & 3;
    const uint32_t w0 = w[q], w1 = w[q + 1], w2 = w[q + 2];
    lo = __builtin_amdgcn_alignbyte(w1, w0, o);
    hi = __builtin_amdgcn_alignbyte(w2, w1, o);
}

// The thread id through an opaque move: the per-lane index arithmetic
// derived from it is redone for each chunk of a workgroup's loop instead of
// being hoisted out and held in registers across it (which left room for 4
// workgroups per CU instead of the 7 the LDS allows).
__device__ __forceinline__ int opaque_tid() {
    int v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "v"((int)threadIdx.x));
    __builtin_assume(v >= 0 && v < kThreads);
    return v;
}

__device__ void decode_chunk(const uint8_t* __restrict__ in, const ChunkGeom& g, const int64_t* __restrict__ offsets,
                             uint8_t* __restrict__ bp, int32_t* __restrict__ err, int64_t in_len, int64_t id,
                             DecodeLds& S) {
    uint32_t* s_c = S.c;
    uint64_t* s_b0 = S.b0;
    uint8_t* s_b1 = S.b1;
    uint8_t* s_b2 = S.b2;
    int* s_scan = S.scan;
    const int tid = opaque_tid();
    const ChunkPos cp = chunk_pos(g, id);
    const int osize = cp.osize;
    const uint16_t* sizes = reinterpret_cast<const uint16_t*>(in + 8);
    const int csize = sizes[id];
    const int64_t start = 8 + 2 * g.nchunks + offsets[id];
    uint64_t* dst0 = reinterpret_cast<uint64_t*>(bp + cp.base);
    if (*reinterpret_cast<const uint64_t*>(in) != (uint64_t)g.host_size || start + csize + 3 > in_len) {
        if (tid == 0) atomicOr(err, 8);
        for (int i = tid; i < (osize >> 3); i += kThreads) dst0[i] = 0ull;
        return;
    }
    const uint8_t* cbg = in + start;
    uint64_t* dst = reinterpret_cast<uint64_t*>(bp + cp.base);
    const int nwords = osize >> 3;
    if (csize > osize || csize < 2) {
        if (tid == 0) atomicOr(err, 1);
        for (int i = tid; i < nwords; i += kThreads) dst[i] = 0ull;
        return;
    }
    if (csize == 2 && osize > 2) {  // an empty chunk, its footer alone: zeros, no staging
        const int orig = (int)cbg[0] | ((int)cbg[1] << 8);
        if (orig != osize && tid == 0) atomicOr(err, 2);
        for (int i = tid; i < nwords; i += kThreads) dst[i] = 0ull;
        return;
    }
    const int shift = (int)((uintptr_t)cbg & 3);
    const uint32_t* aw = reinterpret_cast<const uint32_t*>(cbg - shift);
    const int cw = (shift + csize + 3) >> 2;
    {  // all loads of a lane in flight before the first LDS write
        constexpr int kPer = (kCS / 4 + 2 + kThreads - 1) / kThreads;
        uint32_t t[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = tid + k * kThreads;
            t[k] = i < cw ? aw[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = tid + k * kThreads;
            if (i < cw) s_c[i] = t[k];
        }
    }
    __syncthreads();
    const uint8_t* cb = reinterpret_cast<const uint8_t*>(s_c) + shift;
    if (csize == osize) {  // stored raw
        for (int i = tid; i < nwords; i += kThreads) {
            uint32_t lo, hi;
            bytes8(s_c, shift + 8 * i, lo, hi);
            dst[i] = (uint64_t)hi << 32 | lo;
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
    // Output word wi holds bytes 8 sub .. 8 sub + 7 of lane t = wi / 8's 64
    // (sub = wi mod 8); its nonzero bytes are the next popc(mask) payload
    // bytes from e0[t] plus the set bits of f0[t] below byte sub.  Lanes
    // take consecutive words, so stores are contiguous and neighbouring
    // lanes read neighbouring payload bytes.
    S.e0[tid] = e0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int wi = tid + k * kThreads;
        if (wi >= nwords) break;
        const int t = wi >> 3, sub = wi & 7;
        const uint64_t f = s_b0[t];
        const uint32_t m = (uint32_t)(f >> (8 * sub)) & 0xFFu;
        uint64_t v = 0;
        if (m != 0 && !bad) {  // consistent sizes keep every read inside the payload (n0 = d1 < csize)
            uint32_t lo, hi;
            bytes8(s_c, shift + S.e0[t] + __popcll(f & ((1ull << (8 * sub)) - 1ull)), lo, hi);
            const uint64_t sel = S.sel[m];
            v = (uint64_t)__builtin_amdgcn_perm(hi, lo, (uint32_t)(sel >> 32)) << 32 |
                __builtin_amdgcn_perm(hi, lo, (uint32_t)sel);
        }
        dst[wi] = v;
    }
}

template <bool WINDOW>
__global__ __launch_bounds__(kThreads) void decode_kernel(const uint8_t* __restrict__ in, ChunkGeom g,
                                                          const int64_t* __restrict__ offsets,
                                                          uint8_t* __restrict__ bp,
                                                          const int32_t* __restrict__ begin,
                                                          const int32_t* __restrict__ end,
                                                          uint64_t* __restrict__ bytes_read,
                                                          int32_t* __restrict__ err, int64_t in_len) {
    __shared__ DecodeLds S;
    init_selectors(S);
    if (!WINDOW) {
        decode_chunk(in, g, offsets, bp, err, in_len, blockIdx.x, S);
        return;
    }
    const Window w = window_ranges(g, begin, end);
    if (blockIdx.x == 0 && threadIdx.x == 0 && bytes_read)
        *bytes_read += window_bytes(w, offsets, reinterpret_cast<const uint16_t*>(in + 8));
    for (int64_t k = blockIdx.x; k < w.count; k += gridDim.x) {
        decode_chunk(in, g, offsets, bp, err, in_len, window_chunk(w, k), S);
        __syncthreads();  // the next chunk reuses the LDS
    }
}

// ------------------------------------------------- batched retrieval --
// decode workspace of a job: error word (64-byte slot) | offsets
struct DecodeWs {
    int32_t* err;
    int64_t* offsets;
};
__device__ __forceinline__ DecodeWs decode_ws(const RetrJob& J) {
    DecodeWs w;
    w.err = reinterpret_cast<int32_t*>(J.dws);
    w.offsets = reinterpret_cast<int64_t*>(J.dws + 64);
    return w;
}

// One workgroup per job: on the field's first round the error words and the
// payload offsets; then, every round, each level's first live plane of the
// window (state[10..13]): the decode, the inverse shuffle and the
// reconstruction take the planes from there on.
__global__ __launch_bounds__(1024) void retr_prep_kernel(const RetrJob* __restrict__ jobs) {
    const RetrJob& J = jobs[blockIdx.x];
    const DecodeWs w = decode_ws(J);
    const uint16_t* sizes = reinterpret_cast<const uint16_t*>(J.stream + 8);
    if (!J.offsets_ready) {
        if (threadIdx.x < 16) reinterpret_cast<uint32_t*>(J.dws)[threadIdx.x] = 0u;
        size_scan(sizes, J.L.nchunks, w.offsets);
        __syncthreads();  // the offsets, written across the workgroup
    }
    if (threadIdx.x < 64) {
        int live[4];
        live_begins(J.cg, w.offsets, sizes, J.state, J.state + 4, live);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int lr = 0; lr < 4; ++lr) J.state[10 + lr] = live[lr];
        }
    }
}

// exclusive prefix of the per-job selected-chunk counts (one workgroup);
// each job's payload bytes go to its counter
__global__ __launch_bounds__(1024) void retr_count_scan_kernel(const RetrJob* __restrict__ jobs, int nj,
                                                               int64_t* __restrict__ prefix) {
    __shared__ int64_t s_w[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (nj + 1023) / 1024;
    const int lo = tid * per, hi = lo + per < nj ? lo + per : nj;
    int64_t sum = 0;
    for (int j = lo; j < hi; ++j) {
        const RetrJob& J = jobs[j];
        // the chunks of the live planes are decoded; the whole window is counted as read
        sum += window_ranges(J.cg, J.state + 10, J.state + 4).count;
        if (J.bytes_read)
            *J.bytes_read += window_bytes(window_ranges(J.cg, J.state, J.state + 4), decode_ws(J).offsets,
                                          reinterpret_cast<const uint16_t*>(J.stream + 8));
    }
    int64_t x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    int64_t run = x - sum;
    for (int w = 0; w < wave; ++w) run += s_w[w];
    for (int j = lo; j < hi; ++j) {
        prefix[j] = run;
        const RetrJob& J = jobs[j];
        run += window_ranges(J.cg, J.state + 10, J.state + 4).count;
    }
    if (tid == 1023) prefix[nj] = run;
}

// the selected chunks of every job, as one sequence walked by a fixed grid
__global__ __launch_bounds__(kThreads) void retr_decode_kernel(const RetrJob* __restrict__ jobs, int nj,
                                                               const int64_t* __restrict__ prefix) {
    __shared__ DecodeLds S;
    init_selectors(S);
    const int64_t total = prefix[nj];
    for (int64_t i = blockIdx.x; i < total; i += gridDim.x) {
        const int j = find_field(prefix, 0, nj, i);
        const RetrJob& J = jobs[j];
        const DecodeWs w = decode_ws(J);
        const Window win = window_ranges(J.cg, J.state + 10, J.state + 4);  // the live planes
        decode_chunk(J.stream, J.cg, w.offsets, J.bp, w.err, J.stream_len, window_chunk(win, i - prefix[j]), S);
        __syncthreads();  // the next chunk reuses the LDS
    }
}

}  // namespace

// Workgroups of a window decode: as many as stay resident at once (the
// decoder's 22 KB of LDS allows 7 per CU), each walking every grid-th
// selected chunk, so no workgroup waits for a slot behind the others.
unsigned decode_grid(int64_t chunks) {
    const int64_t resident = 7 * (int64_t)device_cus();
    return (unsigned)(chunks < resident ? chunks : resident);
}
}  // namespace pa

using namespace pa;

int pa::retr_decode(const RetrDev& r, hipStream_t st) {
    hipLaunchKernelGGL(retr_prep_kernel, dim3((unsigned)r.nj), dim3(1024), 0, st, r.jobs);
    PA_LAUNCH_CHECK();
    hipLaunchKernelGGL(retr_count_scan_kernel, dim3(1), dim3(1024), 0, st, r.jobs, r.nj, r.p_chunk);
    PA_LAUNCH_CHECK();
    if (r.max_chunks > 0) {
        const unsigned grid = decode_grid(r.max_chunks);
        hipLaunchKernelGGL(retr_decode_kernel, dim3(grid), dim3(kThreads), 0, st, r.jobs, r.nj, r.p_chunk);
        PA_LAUNCH_CHECK();
    }
    return PA_OK;
}

// the error word (an 8-byte slot), then the look-back status words (8 B per chunk)
extern "C" size_t pa_encode_workspace(const pa_layout* L) {
    if (!L) return 0;
    return ((size_t)L->nchunks + 1) * 8;
}

// chunks per encoder workgroup (one per wave64)
constexpr int kEncodeWaves = 4;

extern "C" int pa_lossless_encode(const uint8_t* d_bp, const pa_layout* L, uint8_t* d_out, uint64_t* d_cum,
                                  uint64_t* d_outsize, void* d_ws, size_t ws_bytes, void* stream) {
    if (!L || !d_bp || !d_out || !d_cum || !d_outsize) return PA_ERR_ARG;
    if (!d_ws || ws_bytes < pa_encode_workspace(L) || ((uintptr_t)d_ws & 7) != 0) return PA_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    ChunkGeom g = make_chunk_geom(L);
    uint64_t* ws = static_cast<uint64_t*>(d_ws);
    hipLaunchKernelGGL(encode_clear_kernel, dim3((unsigned)((L->nchunks + 1 + 255) / 256)), dim3(256), 0, st, ws,
                       L->nchunks);
    PA_LAUNCH_CHECK();
    EncodeOut o{d_out, d_cum, d_outsize, ws + 1, reinterpret_cast<int32_t*>(ws)};
    const unsigned eblocks = (unsigned)((L->nchunks + kEncodeWaves - 1) / kEncodeWaves);
    hipLaunchKernelGGL(encode_kernel<kEncodeWaves>, dim3(eblocks), dim3(64 * kEncodeWaves), 0, st, d_bp, g, o);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

extern "C" int pa_compress_merge(uint8_t* d_out, const pa_layout* L, int dtype, const uint64_t* d_outsize,
                                 const uint8_t* d_flags8, const double* d_range, const uint64_t* d_cum,
                                 const void* d_anchors, uint64_t* d_total, void* stream) {
    if (!L || !d_out || !d_outsize || !d_flags8 || !d_cum || !d_anchors || !d_total) return PA_ERR_ARG;
    if (dtype != PA_F32 && dtype != PA_F64) return PA_ERR_ARG;
    const int64_t ab = L->n4 * dtype;
    const int blocks = (int)batch_merge_blocks(L, dtype);
    hipLaunchKernelGGL(merge_kernel, dim3(blocks), dim3(256), 0, as_stream(stream), d_out, d_outsize, d_flags8,
                       d_range, d_cum, static_cast<const uint8_t*>(d_anchors), ab, d_total);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

// workspace: int32 error word (64-byte slot) | int64 offsets[nchunks]
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
        hipLaunchKernelGGL(size_scan_kernel, dim3(1), dim3(1024), 0, st,
                           reinterpret_cast<const uint16_t*>(d_in + 8), L->nchunks, offsets,
                           reinterpret_cast<uint64_t*>(d_ws));
        PA_LAUNCH_CHECK();
    }
    if (window) {
        const unsigned grid = decode_grid(L->nchunks);
        hipLaunchKernelGGL(decode_kernel<true>, dim3(grid), dim3(kThreads), 0, st, d_in, g, offsets, d_bp, d_begin,
                           d_end, d_bytes_read, err, in_len);
    } else {
        hipLaunchKernelGGL(decode_kernel<false>, dim3((unsigned)L->nchunks), dim3(kThreads), 0, st, d_in, g, offsets,
                           d_bp, nullptr, nullptr, nullptr, err, in_len);
    }
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

// ------------------------------------------- allocation-free decode --
// pa_lossless_decode with its workspace (error slot + chunk offsets) in a
// file-scope device array instead of caller memory.  Host code never passes
// `scratch` itself to a kernel: its device address, different on every GPU,
// comes from hipGetSymbolAddress once per device.
namespace {
constexpr size_t kNoAllocScratch = (size_t)1 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratch];

std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// the device address of this GPU's scratch (nullptr on failure)
void* scratch_on_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// copies the decoder's error word out of the scratch
__global__ void publish_error_kernel(const int32_t* __restrict__ err, int32_t* __restrict__ out) {
    if (threadIdx.x == 0) *out = *err;
}
}  // namespace

extern "C" size_t pa_noalloc_scratch_bytes(void) { return kNoAllocScratch; }

extern "C" int pa_lossless_decode_noalloc(const uint8_t* d_in, int64_t in_len, const pa_layout* L, uint8_t* d_bp,
                                          int32_t* d_err, void* stream) {
    if (!L || !d_in || !d_bp || !d_err) return PA_ERR_ARG;
    if (pa_decode_workspace(L) > kNoAllocScratch) return PA_ERR_WORKSPACE;  // before any HIP call
    void* ws = scratch_on_current_device();
    if (!ws) return PA_ERR_HIP;
    hipStream_t st = as_stream(stream);
    const int rc = decode_common(d_in, in_len, L, nullptr, nullptr, d_bp, nullptr, false, false, ws,
                                 kNoAllocScratch, st);
    if (rc != PA_OK) return rc;
    hipLaunchKernelGGL(publish_error_kernel, dim3(1), dim3(64), 0, st, static_cast<const int32_t*>(ws), d_err);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

int64_t pa::batch_merge_blocks(const pa_layout* L, int dtype) {
    const int64_t ab = L->n4 * dtype;
    int64_t blocks = (ab / 4 + 255) / 256;
    return blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
}

template <int kW>
static void launch_batch_encode(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL(batch_encode_kernel<kW>, dim3((unsigned)(b.n_chunk / kW)), dim3(64 * kW), 0, st, b.fields,
                       b.p_chunk, b.nf);
}

// the encoder launch (its status words cleared by the front launch of the run)
static int batch_encode_only(const BatchDev& b, hipStream_t st) {
    if (b.n_chunk <= 0) return PA_OK;
    switch (b.encode_waves) {
        case 2: launch_batch_encode<2>(b, st); break;
        case 8: launch_batch_encode<8>(b, st); break;
        default: launch_batch_encode<4>(b, st); break;
    }
    PA_LAUNCH_CHECK();
    return PA_OK;
}

int pa::batch_encode(const BatchDev& b, hipStream_t st) {
    int e = batch_encode_only(b, st);
    if (e != PA_OK) return e;
    hipLaunchKernelGGL(batch_merge_kernel, dim3((unsigned)b.n_merge), dim3(256), 0, st, b.fields, b.p_merge, b.nf);
    PA_LAUNCH_CHECK();
    return PA_OK;
}

int pa::batch_encode_chunks(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL(batch_encode_clear_kernel, dim3((unsigned)((b.n_chunk + b.nf + 255) / 256)), dim3(256), 0, st,
                       b.fields, b.p_chunk, b.nf, b.n_chunk);
    PA_LAUNCH_CHECK();
    return batch_encode_only(b, st);
}
