/*
 * pipeline.c -- whole compress / decompress / progressive-round flows of the
 * CPU oracle (test infrastructure only; see prism_oracle.h).
 * Follows compressor.cu:120-226 and prism.cu:9-151.
 */
#include "prism_oracle.h"

#include <stdlib.h>
#include <string.h>

int64_t po_compress_bound(uint32_t nx, uint32_t ny, uint32_t nz, int dtype_bytes) {
    po_layout L;
    po_layout_init(&L, nx, ny, nz);
    return po_stream_bound(&L) + L.trailer_fixed + L.n4 * dtype_bytes;
}

/* compress_merge, compressor.cu:139-151 */
static int64_t write_trailer(uint8_t* out, int64_t p, const uint8_t flags[8], double range,
                             const uint64_t cum[128], const void* anchors, int64_t anchor_bytes) {
    memcpy(out + p, flags, 8); /* bool use_md[4], bool reverse[4] */
    memcpy(out + p + 8, &range, 8);
    memcpy(out + p + 16, cum, 128 * 8);
    memcpy(out + p + 16 + 1024, anchors, (size_t)anchor_bytes);
    return p + 1040 + anchor_bytes;
}

int po_parse_trailer(const uint8_t* in, int64_t len, const po_layout* L, int dtype_bytes,
                     uint8_t flags[8], double* range, uint64_t cum[128], int64_t* payload_len) {
    int64_t tsize = L->trailer_fixed + L->n4 * dtype_bytes;
    if (len < tsize + 8 + 2 * L->nchunks) return -1;
    int64_t p = len - tsize;
    memcpy(flags, in + p, 8);
    memcpy(range, in + p + 8, 8);
    memcpy(cum, in + p + 16, 1024);
    *payload_len = p;
    return 0;
}

#define DEF_COMPRESS(SFX, T)                                                                     \
    int64_t po_compress_##SFX(const T* data, uint32_t nx, uint32_t ny, uint32_t nz,              \
                              double eb_or_rel, int abs_mode, int mode, uint8_t* out,            \
                              int64_t out_cap, uint8_t flags_out[8], int nthreads) {             \
        po_layout L;                                                                             \
        po_layout_init(&L, nx, ny, nz);                                                          \
        if (out_cap < po_compress_bound(nx, ny, nz, (int)sizeof(T))) return -1;                 \
        double range = 1.0, eb, rel_eb = 0.0;                                                    \
        if (abs_mode) {                                                                          \
            eb = eb_or_rel;                                                                      \
        } else {                                                                                 \
            double mm[2];                                                                        \
            po_extrema_##SFX(data, L.n, mm);                                                     \
            range = mm[1] - mm[0];                                                               \
            rel_eb = eb_or_rel;                                                                  \
            eb = rel_eb * range;                                                                 \
        }                                                                                        \
        double scores[15];                                                                       \
        uint8_t flags[8];                                                                        \
        po_autotune_##SFX(data, &L, rel_eb, scores, flags);                                      \
        int32_t* qc = (int32_t*)malloc(sizeof(int32_t) * (size_t)L.code_len);                   \
        T* anchors = (T*)malloc(sizeof(T) * (size_t)L.n4);                                       \
        uint8_t* bp = (uint8_t*)malloc((size_t)L.host_size);                                     \
        po_spline_construct_##SFX(data, &L, eb, rel_eb, flags, qc, anchors, nthreads);           \
        po_bitshuffle(qc, &L, mode, bp);                                                         \
        uint64_t cum[128];                                                                       \
        int64_t plen = po_lossless_encode(bp, &L, out, cum);                                     \
        int64_t total = write_trailer(out, plen, flags, range, cum, anchors,                     \
                                      L.n4 * (int64_t)sizeof(T));                                \
        memcpy(flags_out, flags, 8);                                                             \
        free(qc);                                                                                \
        free(anchors);                                                                           \
        free(bp);                                                                                \
        return total;                                                                            \
    }

DEF_COMPRESS(f32, float)
DEF_COMPRESS(f64, double)

#define DEF_DECOMPRESS(SFX, T)                                                                   \
    int po_decompress_##SFX(const uint8_t* in, int64_t len, uint32_t nx, uint32_t ny,            \
                            uint32_t nz, double eb_or_rel, int abs_mode, T* out, int nthreads) { \
        po_layout L;                                                                             \
        po_layout_init(&L, nx, ny, nz);                                                          \
        uint8_t flags[8];                                                                        \
        double range;                                                                            \
        uint64_t cum[128];                                                                       \
        int64_t plen;                                                                            \
        if (po_parse_trailer(in, len, &L, (int)sizeof(T), flags, &range, cum, &plen)) return -1; \
        double eb = abs_mode ? eb_or_rel : eb_or_rel * range;                                    \
        double rel_eb = abs_mode ? 0.0 : eb_or_rel;                                              \
        uint8_t* bp = (uint8_t*)calloc((size_t)L.host_size, 1);                                  \
        int32_t* qc = (int32_t*)calloc((size_t)L.code_len, sizeof(int32_t));                     \
        int rc = po_lossless_decode(in, plen, &L, bp);                                           \
        if (rc == 0) {                                                                           \
            po_inverse_bitshuffle(bp, &L, 1, qc);                                                \
            po_spline_reconstruct_##SFX(qc, (const T*)(in + plen + 1040), &L, eb, rel_eb, flags, \
                                        4, out, nthreads);                                       \
        }                                                                                        \
        free(bp);                                                                                \
        free(qc);                                                                                \
        return rc;                                                                               \
    }

DEF_DECOMPRESS(f32, float)
DEF_DECOMPRESS(f64, double)

#define DEF_PROGRESSIVE(SFX, T)                                                                  \
    int po_progressive_round_##SFX(const uint8_t* in, int64_t len, uint32_t nx, uint32_t ny,     \
                                   uint32_t nz, double eb_or_rel, int abs_mode, double target,   \
                                   int first, int corrected, int32_t begin[4], int32_t end[4],   \
                                   const T* old, T* out, int64_t* bytes_read, int nthreads) {    \
        po_layout L;                                                                             \
        po_layout_init(&L, nx, ny, nz);                                                          \
        uint8_t flags[8];                                                                        \
        double range;                                                                            \
        uint64_t cum[128];                                                                       \
        int64_t plen;                                                                            \
        if (po_parse_trailer(in, len, &L, (int)sizeof(T), flags, &range, cum, &plen)) return -1; \
        double eb = abs_mode ? eb_or_rel : eb_or_rel * range;                                    \
        double rel_eb = abs_mode ? 0.0 : eb_or_rel;                                              \
        double target_abs = abs_mode ? target : target * range;                                  \
        int maxlevel = po_find_strategy(cum, begin, end, eb, target_abs, NULL);                  \
        uint8_t* bp = (uint8_t*)calloc((size_t)L.host_size, 1);                                  \
        int32_t* qc = (int32_t*)calloc((size_t)L.code_len, sizeof(int32_t));                     \
        int64_t nb = po_lossless_decode_window(in, plen, &L, begin, end, corrected, bp);         \
        if (nb < 0) {                                                                            \
            free(bp);                                                                            \
            free(qc);                                                                            \
            return -2;                                                                           \
        }                                                                                        \
        if (bytes_read) *bytes_read = nb;                                                        \
        po_inverse_bitshuffle_window(bp, &L, 0, begin, end, qc);                                 \
        if (first)                                                                               \
            po_spline_reconstruct_##SFX(qc, (const T*)(in + plen + 1040), &L, eb, rel_eb, flags, \
                                        maxlevel, out, nthreads);                                \
        else                                                                                     \
            po_spline_delta_reconstruct_##SFX(qc, &L, eb, rel_eb, flags, maxlevel, old, out,     \
                                              nthreads);                                         \
        free(bp);                                                                                \
        free(qc);                                                                                \
        return maxlevel;                                                                         \
    }

DEF_PROGRESSIVE(f32, float)
DEF_PROGRESSIVE(f64, double)
