/*
 * prism_oracle.h -- CPU oracle for the PRISM compress / progressive-retrieve path.
 *
 * TEST INFRASTRUCTURE. This is a plain-C restatement of the reference
 * algorithm, written to be checked against closed forms and invariants.
 * It is loaded only by the test suite, by the smoke check and by the CPU
 * baseline timing in bench.py.  The product library (prism_amd) never links,
 * loads or calls anything in this directory.
 *
 * Parity status of each function is stated next to it: "pinned" means a test
 * checks it against a closed form, an invariant or an independent emulation
 * of the reference's device code; "parity unpinned" means only the
 * reference's own run could pin the exact bits (see the notes there).  The
 * reference holds no test vectors, fixtures or known-answer tests for this
 * path, and it cannot be built here (CUDA only), so every pin below is a
 * closed form, an invariant, an emulation or a brute force.
 *
 * Conventions (all citations are paths inside the reference tree):
 *   - arrays are T[nz][ny][nx], x fastest;
 *   - level_id 0..3 = unit 1,2,4,8 (level_id 4 = anchors, stride 16);
 *   - "lr" = 3 - level_id, the index used by begin/end and the size table;
 *   - all floating point is evaluated in the order written, with no FMA
 *     contraction (built with -ffp-contract=off).
 */
#ifndef PRISM_ORACLE_H
#define PRISM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PO_CS 16384 /* codec chunk size, include/encoder/lossless.hpp:4 */

/* Field geometry and every derived size used by the stream layout.
 * Follows InterpolationPredictor.cu:969-995 (level grids, device prefix),
 * bitplane.cu:10-53 (device strides) and io.hpp:104-132 + io.cu:117-127
 * (host strides used to cut codec chunks). */
typedef struct po_layout {
    uint32_t nx, ny, nz;
    int64_t n;          /* nx*ny*nz */
    int64_t n4;         /* anchors: ceil(nx/16)*ceil(ny/16)*ceil(nz/16) */
    int64_t lx[5], lxy[5]; /* level-grid leaps: level 0 = full grid ... 4 = anchors */
    int64_t cnt[4];     /* cnt[i] = #points of the stride-2^(i+1) grid */
    int64_t prefix[5];  /* absolute level-major start of level i (device pads), prefix[4]=0 */
    int64_t dev_align;  /* total device padding (incl. level-0 tail) */
    int64_t real_end;   /* absolute end of real codes (start of level-0 phantom tail) */
    int64_t code_len;   /* n + dev_align: absolute length of the code array */
    int64_t dev_len[4];    /* codes in each level segment (device, incl. pads/phantom) */
    int64_t dev_stride[4]; /* s_l: bytes per bit-plane of level l (device) */
    int64_t dev_base[4];   /* A_l: start of level l in the bit-plane buffer */
    int64_t dev_size;      /* sum 32*s_l */
    int64_t host_stride[4];/* e_l: bytes per bit-plane of level l (host, codec chunking) */
    int64_t host_base[4];  /* E_l */
    int64_t host_size;     /* sum 32*e_l: bit-plane buffer size and stream 'insize' */
    int64_t chunks_per_plane[4]; /* ceil(e_l / CS) */
    int64_t nchunks;
    int64_t trailer_fixed; /* 1040 bytes before the anchors */
} po_layout;

/* Pinned: sizes of eight shapes against tests/golden/layout_sizes.json (the
 * sizes restated from io.hpp:104-132 and bitplane.cu:10-53 by hand), host
 * strides >= device strides, level counts = lattice point counts. */
void po_layout_init(po_layout* L, uint32_t nx, uint32_t ny, uint32_t nz);

/* Level-major position of grid point (gx,gy,gz) in the code array.
 * InterpolationPredictor.cu:2070-2083.  Pinned: equals the point's rank in
 * raster order among the points of its level (brute force over all points). */
int64_t po_lm_index(const po_layout* L, uint32_t gx, uint32_t gy, uint32_t gz);

/* ---- error bound schedule, InterpolationPredictor.cu:2248-2271, 1825-1839 ----
 * Pinned: alpha at its breakpoints and midpoints (closed form); per-level
 * bound eb / min(alpha^k, 4) is met and more than half used on noise. */
double po_alpha(double rel_eb);
/* per-level (eb_r, ebx2) in T precision for unit u; is_double selects T */
void po_level_eb_f32(float eb_r, float ebx2, double alpha, double beta, int unit,
                     float* cur_eb_r, float* cur_ebx2);
void po_level_eb_f64(double eb_r, double ebx2, double alpha, double beta, int unit,
                     double* cur_eb_r, double* cur_ebx2);

/* ---- extrema, compressor.cu:228-302 ----
 * Pinned: the trailer's range equals numpy's max - min (pipeline test). */
void po_extrema_f32(const float* in, int64_t n, double out_minmax[2]);
void po_extrema_f64(const double* in, int64_t n, double out_minmax[2]);

/* ---- predictor: interpolation kernel, InterpolationPredictor.cu:2137-2173 ----
 * flags[0..3] = use_md[level_id], flags[4..7] = reverse[level_id].
 * qc must hold L->code_len int32 and is fully written (pads and the anchor
 * slots are set to 0).  anchors holds L->n4 values.
 * Pinned (construct, reconstruct, delta): linear fields give zero codes and
 * exact reconstruction under every flag setting; single-code impulse
 * responses equal hand-derived stencil weights (1-D, face, cube); the bound
 * holds; zero delta codes leave old unchanged.  Parity unpinned: whether
 * nvcc contracts any of these products into FMAs in the reference build
 * (this oracle and the HIP library evaluate without contraction). */
void po_spline_construct_f32(const float* data, const po_layout* L, double eb, double rel_eb,
                             const uint8_t flags[8], int32_t* qc, float* anchors, int nthreads);
void po_spline_construct_f64(const double* data, const po_layout* L, double eb, double rel_eb,
                             const uint8_t flags[8], int32_t* qc, double* anchors, int nthreads);

/* full reconstruction, interpolation_reverse InterpolationPredictor.cu:2177-2207 */
void po_spline_reconstruct_f32(const int32_t* qc, const float* anchors, const po_layout* L,
                               double eb, double rel_eb, const uint8_t flags[8], int maxlevel,
                               float* out, int nthreads);
void po_spline_reconstruct_f64(const int32_t* qc, const double* anchors, const po_layout* L,
                               double eb, double rel_eb, const uint8_t flags[8], int maxlevel,
                               double* out, int nthreads);
/* delta reconstruction + old, InterpolationPredictor.cu:2209-2239, 1123-1147 */
void po_spline_delta_reconstruct_f32(const int32_t* qc, const po_layout* L, double eb,
                                     double rel_eb, const uint8_t flags[8], int maxlevel,
                                     const float* old, float* out, int nthreads);
void po_spline_delta_reconstruct_f64(const int32_t* qc, const po_layout* L, double eb,
                                     double rel_eb, const uint8_t flags[8], int maxlevel,
                                     const double* old, double* out, int nthreads);

/* ---- auto-tuner, InterpolationPredictor.cu:880-967 + 2273-2355 ----
 * scores[15] receive the candidate error sums in the reference's index order.
 * Returns 0 if tuning ran, 1 if the shape leaves it undefined (flags all 0).
 * Pinned: fields varying along x only (z only) select x-first (z-first)
 * order at every level; shapes with an axis < 17 give all-false flags.
 * Parity unpinned: the exact score bits, which the reference sums with float
 * atomics in no fixed order (here: a fixed-order double sum). */
int po_autotune_f32(const float* data, const po_layout* L, double rel_eb,
                    double scores[15], uint8_t flags[8]);
int po_autotune_f64(const double* data, const po_layout* L, double rel_eb,
                    double scores[15], uint8_t flags[8]);
/* Host flag pick from the 15 scores, InterpolationPredictor.cu:2306-2355.
 * Pinned: strict/non-strict comparisons checked on tie cases. */
void po_pick_flags(const double scores[15], uint8_t flags[8]);

/* ---- bit-plane shuffle, bitplane.cu:55-253 (closed form), 256-601 ----
 * mode 0 = NB (negabinary, progressive), 1 = SM (sign-magnitude variant).
 * bp holds L->host_size bytes and is fully written.
 * Pinned: equals a lane-by-lane emulation of the reference's warp network
 * (bitplane.cu:90-132, its inverse 292-318); NB/SM are involutions; single
 * codes land on the predicted bytes and bits; windowed inverses add up.
 * Parity unpinned: the phantom-tail bytes past the last real code, which the
 * reference fills from reads past its code array (here: zero codes). */
void po_bitshuffle(const int32_t* qc, const po_layout* L, int mode, uint8_t* bp);
/* inverse; writes qc[n4 .. real_end) (other entries untouched) */
void po_inverse_bitshuffle(const uint8_t* bp, const po_layout* L, int mode, int32_t* qc);
/* plane-window inverse: only planes [begin[lr], end[lr]) of each level */
void po_inverse_bitshuffle_window(const uint8_t* bp, const po_layout* L, int mode,
                                  const int32_t begin[4], const int32_t end[4], int32_t* qc);

/* ---- chunk codec, RZE.hpp:10-167 ----
 * Encodes one chunk of csize bytes; returns the encoded length, or -1 when
 * the chunk must be stored raw. out needs PO_CS bytes.
 * Pinned: a hand-encoded example byte for byte, the all-zero chunk (footer
 * only), raw storage of incompressible chunks, size limits, round trips. */
int po_rze_encode_chunk(const uint8_t* in, int csize, uint8_t* out);
/* Decodes one encoded chunk of clen bytes; returns decoded length or -1. */
int po_rze_decode_chunk(const uint8_t* in, int clen, uint8_t* out);

/* ---- lossless stream, zbpc.cu:155-263 + 476-520 ----
 * out needs po_stream_bound() bytes.  Writes header, size table, payloads
 * and the zero pad to 8; returns the padded length.  cum[4*32] receives the
 * inclusive payload size through the last chunk of each (lr, plane).
 * Pinned: stream round trip, size table and cum table consistent with the
 * payload, window decodes select exactly the covering chunks, and reading
 * with the reference's selection down to the bound reads each byte once. */
int64_t po_stream_bound(const po_layout* L);
int64_t po_lossless_encode(const uint8_t* bp, const po_layout* L, uint8_t* out, uint64_t cum[128]);
/* full decode, zbpc.cu:265-365: returns 0 on success, <0 on a malformed stream */
int po_lossless_decode(const uint8_t* in, int64_t in_len, const po_layout* L, uint8_t* bp);
/* progressive decode, zbpc.cu:368-473.  corrected==0 decodes the host
 * chunks [begin*cb, end*cb) of each level (the reference's selection);
 * corrected==1 decodes every chunk that holds a byte of a device plane in
 * [begin, end).  Chunks not selected leave bp untouched.  Returns the
 * number of payload bytes read, or <0 on a malformed stream. */
int64_t po_lossless_decode_window(const uint8_t* in, int64_t in_len, const po_layout* L,
                                  const int32_t begin[4], const int32_t end[4], int corrected,
                                  uint8_t* bp);

/* ---- strategy search, dataloader.cu:21-208 ----
 * Updates begin/end in place (begin <- end, end <- max(end, choice)) and
 * returns maxlevel.  uses_search is set to 1 when the search ran, 0 when
 * the 'load all' path was taken.
 * Pinned: costs equal the double-precision evaluation of the reference's
 * formula; the search equals a literal emulation of its 33 x 1024 tree
 * reductions with their tie-breaks; update rules checked case by case.
 * Parity unpinned: cost boundary cases where CUDA's pow/ceil could round an
 * integer cost differently (exact integer arithmetic here). */
int po_find_strategy(const uint64_t cum[128], int32_t begin[4], int32_t end[4], double eb,
                     double target_error, int* uses_search);
/* raw search result for a given target cost (no begin/end update) */
void po_strategy_search(const uint64_t cum[128], uint64_t target_cost, int32_t choice[4],
                        uint64_t* best_value);
uint64_t po_strategy_cost(int lr, int b);

/* ---- whole pipelines (compressor.cu:120-190, prism.cu:9-151) ----
 * Pinned: stream layout fields (header size, flags, range, anchors = raw
 * values at the 16-multiples, pad), bound met by decompress and by every
 * progressive round (tests/test_oracle_pipeline.py). */
/* compress: returns total stream length (payload + pad + trailer), writes
 * flags used into flags_out.  mode 0 = NB (-prog), 1 = SM.  abs_mode != 0
 * treats eb_or_rel as an absolute bound (range stored as 1.0). */
int64_t po_compress_f32(const float* data, uint32_t nx, uint32_t ny, uint32_t nz, double eb_or_rel,
                        int abs_mode, int mode, uint8_t* out, int64_t out_cap, uint8_t flags_out[8],
                        int nthreads);
int64_t po_compress_f64(const double* data, uint32_t nx, uint32_t ny, uint32_t nz, double eb_or_rel,
                        int abs_mode, int mode, uint8_t* out, int64_t out_cap, uint8_t flags_out[8],
                        int nthreads);
int64_t po_compress_bound(uint32_t nx, uint32_t ny, uint32_t nz, int dtype_bytes);
/* non-progressive decompress (SM stream) */
int po_decompress_f32(const uint8_t* in, int64_t len, uint32_t nx, uint32_t ny, uint32_t nz,
                      double eb_or_rel, int abs_mode, float* out, int nthreads);
int po_decompress_f64(const uint8_t* in, int64_t len, uint32_t nx, uint32_t ny, uint32_t nz,
                      double eb_or_rel, int abs_mode, double* out, int nthreads);

/* One progressive round (compressor.cu:166-190).  state holds begin/end
 * (both start at 0).  first != 0 runs the full reconstruction into out,
 * otherwise out = old + delta.  Returns maxlevel, or <0 on error.
 * bytes_read (may be NULL) receives the payload bytes decoded. */
int po_progressive_round_f32(const uint8_t* in, int64_t len, uint32_t nx, uint32_t ny, uint32_t nz,
                             double eb_or_rel, int abs_mode, double target, int first,
                             int corrected, int32_t begin[4], int32_t end[4], const float* old,
                             float* out, int64_t* bytes_read, int nthreads);
int po_progressive_round_f64(const uint8_t* in, int64_t len, uint32_t nx, uint32_t ny, uint32_t nz,
                             double eb_or_rel, int abs_mode, double target, int first,
                             int corrected, int32_t begin[4], int32_t end[4], const double* old,
                             double* out, int64_t* bytes_read, int nthreads);

/* trailer parse helper (compressor.cu:192-210); returns 0 on success */
int po_parse_trailer(const uint8_t* in, int64_t len, const po_layout* L, int dtype_bytes,
                     uint8_t flags[8], double* range, uint64_t cum[128], int64_t* payload_len);

#ifdef __cplusplus
}
#endif
#endif
