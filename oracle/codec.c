/*
 * codec.c -- zero/repetition-elimination chunk codec and the lossless stream
 * of the CPU oracle (test infrastructure only; see prism_oracle.h).
 *
 * Chunk format (RZE.hpp:10-86 with T = byte, d_zero_elimination.hpp:5-73,
 * d_repetition_elimination.hpp:6-146 and 243-370):
 *   [non-zero bytes of the chunk]            zero elimination, bitmap B0 (2048 B)
 *   [bytes of B0 that differ from the previous byte (first vs 0)]   -> B1 (256 B)
 *   [bytes of B1 that differ ...]                                    -> B2 (32 B)
 *   [bytes of B2 that differ ...]                                    -> B3 (4 B)
 *   [B3]
 *   [original size, 2 bytes little endian]
 * All bitmaps are LSB-first within a byte.  An all-zero chunk is the 2-byte
 * footer only.  The encoder gives up (chunk stored raw) when a stage would
 * pass CS-2 bytes (RZE.hpp:34,43,50,58,66) or the result is not smaller than
 * the chunk (zbpc.cu:230).
 */
#include "prism_oracle.h"

#include <stdlib.h>
#include <string.h>

static int popcount_bytes(const uint8_t* b, int nbits) {
    int c = 0;
    for (int i = 0; i < nbits; ++i) c += (b[i >> 3] >> (i & 7)) & 1;
    return c;
}

/* repetition elimination of a[n]; returns the number of kept bytes or -1
 * when it exceeds cap (checked before anything is written, like the
 * reference's __syncthreads_or test) */
static int re_encode(const uint8_t* a, int n, uint8_t* bm, uint8_t* out, int cap) {
    int cnt = 0;
    uint8_t prev = 0;
    memset(bm, 0, (size_t)(n / 8));
    for (int i = 0; i < n; ++i) {
        if (a[i] != prev) {
            bm[i >> 3] |= (uint8_t)(1u << (i & 7));
            cnt++;
        }
        prev = a[i];
    }
    if (cnt > cap) return -1;
    int k = 0;
    for (int i = 0; i < n; ++i)
        if ((bm[i >> 3] >> (i & 7)) & 1) out[k++] = a[i];
    return cnt;
}

int po_rze_encode_chunk(const uint8_t* in, int csize, uint8_t* out) {
    const int avail = PO_CS - 2;
    uint8_t bm[2048 + 256 + 32 + 4];
    memset(bm, 0, sizeof(bm));
    int wpos = 0;
    for (int i = 0; i < csize; ++i) {
        if (in[i]) {
            bm[i >> 3] |= (uint8_t)(1u << (i & 7));
            out[wpos++] = in[i];
        }
    }
    if (wpos >= avail) return -1;
    if (wpos != 0) {
        int c = re_encode(bm, 2048, bm + 2048, out + wpos, avail - wpos);
        if (c < 0) return -1;
        wpos += c;
        c = re_encode(bm + 2048, 256, bm + 2304, out + wpos, avail - wpos);
        if (c < 0) return -1;
        wpos += c;
        c = re_encode(bm + 2304, 32, bm + 2336, out + wpos, avail - wpos);
        if (c < 0) return -1;
        wpos += c;
        if (wpos >= avail - 4) return -1;
        memcpy(out + wpos, bm + 2336, 4);
        wpos += 4;
    }
    out[wpos] = (uint8_t)(csize & 0xFF);
    out[wpos + 1] = (uint8_t)((csize >> 8) & 0xFF);
    return wpos + 2;
}

/* inverse of re_encode: the kept bytes end at rpos; returns their start */
static int re_decode(const uint8_t* in, int rpos, const uint8_t* flags, int n, uint8_t* out) {
    int cnt = popcount_bytes(flags, n);
    if (cnt > rpos) return -1;
    int start = rpos - cnt, k = start;
    uint8_t prev = 0;
    for (int i = 0; i < n; ++i) {
        if ((flags[i >> 3] >> (i & 7)) & 1) prev = in[k++];
        out[i] = prev;
    }
    return start;
}

int po_rze_decode_chunk(const uint8_t* in, int clen, uint8_t* out) {
    if (clen < 2) return -1;
    int rpos = clen - 2;
    int orig = in[rpos] | (in[rpos + 1] << 8);
    if (orig > PO_CS) return -1;
    if (rpos == 0) {
        memset(out, 0, (size_t)orig);
        return orig;
    }
    if (rpos < 4) return -1;
    uint8_t bm[2048 + 256 + 32 + 4];
    rpos -= 4;
    memcpy(bm + 2336, in + rpos, 4);
    rpos = re_decode(in, rpos, bm + 2336, 32, bm + 2304);
    if (rpos < 0) return -1;
    rpos = re_decode(in, rpos, bm + 2304, 256, bm + 2048);
    if (rpos < 0) return -1;
    rpos = re_decode(in, rpos, bm + 2048, 2048, bm);
    if (rpos < 0) return -1;
    int k = 0;
    for (int i = 0; i < orig; ++i) {
        if ((bm[i >> 3] >> (i & 7)) & 1) {
            if (k >= rpos) return -1;
            out[i] = in[k++];
        } else {
            out[i] = 0;
        }
    }
    if (k != rpos) return -1;
    return orig;
}

int64_t po_stream_bound(const po_layout* L) { return 8 + 2 * L->nchunks + L->host_size + 8; }

static inline void put_u16(uint8_t* p, uint16_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
}
static inline uint16_t get_u16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }

/* Chunk order (zbpc.cu:185-206): levels 3..0, planes 0..31, chunk index
 * ascending; chunk (l,b,c) covers host bytes
 * [E_l + b*e_l + c*CS, min(+CS, E_l + (b+1)*e_l)). */
int64_t po_lossless_encode(const uint8_t* bp, const po_layout* L, uint8_t* out, uint64_t cum[128]) {
    int64_t hdr = 8 + 2 * L->nchunks;
    uint8_t* payload = out + hdr;
    int64_t off = 0, id = 0;
    uint8_t tmp[PO_CS];
    for (int l = 3; l >= 0; --l) {
        int64_t e = L->host_stride[l], cb = L->chunks_per_plane[l];
        for (int b = 0; b < 32; ++b) {
            for (int64_t c = 0; c < cb; ++c, ++id) {
                int64_t base = L->host_base[l] + b * e + c * PO_CS;
                int osize = (int)((e - c * PO_CS) < PO_CS ? (e - c * PO_CS) : PO_CS);
                int cs = po_rze_encode_chunk(bp + base, osize, tmp);
                if (cs < 0 || cs >= osize) {
                    cs = osize;
                    memcpy(payload + off, bp + base, (size_t)osize);
                } else {
                    memcpy(payload + off, tmp, (size_t)cs);
                }
                put_u16(out + 8 + 2 * id, (uint16_t)cs);
                off += cs;
                if (c == cb - 1) cum[(3 - l) * 32 + b] = (uint64_t)off;
            }
        }
    }
    uint64_t insize = (uint64_t)L->host_size;
    memcpy(out, &insize, 8);
    int64_t total = hdr + off;
    int64_t pad = (8 - (total % 8)) % 8; /* zbpc.cu:513-516 */
    memset(out + total, 0, (size_t)pad);
    return total + pad;
}

/* payload offsets of all chunks; returns 0 or -1 if the table overruns */
static int chunk_offsets(const uint8_t* in, int64_t payload_end, const po_layout* L, int64_t* offs) {
    int64_t hdr = 8 + 2 * L->nchunks;
    if (payload_end < hdr) return -1;
    int64_t off = hdr;
    for (int64_t id = 0; id < L->nchunks; ++id) {
        offs[id] = off;
        off += get_u16(in + 8 + 2 * id);
    }
    if (off > payload_end) return -1;
    return 0;
}

static int decode_one(const uint8_t* in, int64_t off, int cs, int osize, uint8_t* dst) {
    if (cs > osize || cs < 2) return -1;
    if (cs == osize) {
        memcpy(dst, in + off, (size_t)osize);
        return 0;
    }
    int r = po_rze_decode_chunk(in + off, cs, dst);
    return r == osize ? 0 : -1;
}

int po_lossless_decode(const uint8_t* in, int64_t in_len, const po_layout* L, uint8_t* bp) {
    if (in_len < 8) return -1;
    uint64_t insize;
    memcpy(&insize, in, 8);
    if ((int64_t)insize != L->host_size) return -2;
    int64_t* offs = (int64_t*)malloc(sizeof(int64_t) * (size_t)L->nchunks);
    if (chunk_offsets(in, in_len, L, offs)) { free(offs); return -3; }
    int64_t id = 0;
    int rc = 0;
    for (int l = 3; l >= 0 && rc == 0; --l) {
        int64_t e = L->host_stride[l], cb = L->chunks_per_plane[l];
        for (int b = 0; b < 32 && rc == 0; ++b)
            for (int64_t c = 0; c < cb; ++c, ++id) {
                int64_t base = L->host_base[l] + b * e + c * PO_CS;
                int osize = (int)((e - c * PO_CS) < PO_CS ? (e - c * PO_CS) : PO_CS);
                if (decode_one(in, offs[id], get_u16(in + 8 + 2 * id), osize, bp + base)) { rc = -4; break; }
            }
    }
    free(offs);
    return rc;
}

int64_t po_lossless_decode_window(const uint8_t* in, int64_t in_len, const po_layout* L,
                                  const int32_t begin[4], const int32_t end[4], int corrected,
                                  uint8_t* bp) {
    if (in_len < 8) return -1;
    uint64_t insize;
    memcpy(&insize, in, 8);
    if ((int64_t)insize != L->host_size) return -2;
    int64_t* offs = (int64_t*)malloc(sizeof(int64_t) * (size_t)L->nchunks);
    if (chunk_offsets(in, in_len, L, offs)) { free(offs); return -3; }
    /* device byte range of the requested planes of every level */
    int64_t lo[4], hi[4];
    for (int l = 0; l < 4; ++l) {
        int lr = 3 - l;
        if (begin[lr] < end[lr]) {
            lo[l] = L->dev_base[l] + (int64_t)begin[lr] * L->dev_stride[l];
            hi[l] = L->dev_base[l] + (int64_t)end[lr] * L->dev_stride[l];
        } else {
            lo[l] = hi[l] = 0;
        }
    }
    int64_t id = 0, bytes = 0;
    for (int l = 3; l >= 0; --l) {
        int64_t e = L->host_stride[l], cb = L->chunks_per_plane[l];
        int lr = 3 - l;
        for (int b = 0; b < 32; ++b)
            for (int64_t c = 0; c < cb; ++c, ++id) {
                int64_t base = L->host_base[l] + b * e + c * PO_CS;
                int osize = (int)((e - c * PO_CS) < PO_CS ? (e - c * PO_CS) : PO_CS);
                int sel;
                if (!corrected) {
                    /* zbpc.cu:420: realID in [begin*cb, end*cb) */
                    int64_t real = b * cb + c;
                    sel = real >= (int64_t)begin[lr] * cb && real < (int64_t)end[lr] * cb;
                } else {
                    sel = 0;
                    for (int k = 0; k < 4; ++k)
                        if (lo[k] < hi[k] && base < hi[k] && base + osize > lo[k]) sel = 1;
                }
                if (!sel) continue;
                int cs = get_u16(in + 8 + 2 * id);
                if (decode_one(in, offs[id], cs, osize, bp + base)) { free(offs); return -4; }
                bytes += cs;
            }
    }
    free(offs);
    return bytes;
}
