/*
 * prism_amd.h -- C ABI of the MI355X (gfx950) PRISM compress / progressive
 * retrieve library (libprism_amd.so).
 *
 * One entry point per stage function of the reference's pipeline; each
 * comment names the reference interface it replaces.  All pointers named
 * d_* are device pointers; `stream` is a hipStream_t passed as void* (NULL =
 * the null stream).  Every call is stream-ordered and returns an int status
 * (PA_OK or a negative PA_ERR_*); no call synchronises the host unless its
 * comment says so.  Buffers are owned by the caller: scratch memory comes in
 * through a workspace argument sized by the matching *_workspace() query.
 *
 * Data are T[nz][ny][nx] with x fastest; dtype is PA_F32 or PA_F64 (the
 * reference's prism_dtype values F4 / F8, include/utils/traits.hpp:21-35).
 * flags8 = {use_md[0..3], reverse[0..3]} indexed by level_id (0 = unit 1).
 * begin/end are int32[4] indexed by lr = 3 - level_id, like the reference.
 */
#ifndef PRISM_AMD_H
#define PRISM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PA_OK 0
#define PA_ERR_ARG (-1)       /* invalid argument or unsupported shape */
#define PA_ERR_HIP (-2)       /* a HIP runtime call failed */
#define PA_ERR_WORKSPACE (-3) /* workspace too small */
#define PA_ERR_STREAM (-4)    /* malformed compressed stream */
#define PA_ERR_CAPACITY (-5)  /* output buffer too small */

#define PA_F32 4
#define PA_F64 8
#define PA_MODE_NB 0 /* negabinary bit-planes: progressive streams ("-prog") */
#define PA_MODE_SM 1 /* sign-magnitude bit-planes: plain streams (default) */

/* Geometry and stream layout of one field.  Same content as the sizes the
 * reference derives in Bitplane::h_comput (io.hpp:104-132), d_comput
 * (bitplane.cu:10-53), pre_compute (InterpolationPredictor.cu:969-995) and
 * lossless_encode (zbpc.cu:478-481). */
typedef struct pa_layout {
    uint32_t nx, ny, nz;
    int64_t n, n4;
    int64_t lx[5], lxy[5];
    int64_t cnt[4];
    int64_t prefix[5];
    int64_t dev_align, real_end, code_len;
    int64_t dev_len[4], dev_stride[4], dev_base[4], dev_size;
    int64_t host_stride[4], host_base[4], host_size;
    int64_t chunks_per_plane[4], nchunks;
    int64_t trailer_fixed;
} pa_layout;

/* Bitplane::Bitplane + calculate_aligned_buffer_size (io.hpp:86-92,
 * io.cu:117-127) and Compressor::init sizes (compressor.cu:77-117). Host only. */
int pa_layout_init(pa_layout* L, uint32_t nx, uint32_t ny, uint32_t nz);
/* upper bound of a compressed stream (payload + pad + trailer) in bytes */
int64_t pa_compress_bound(const pa_layout* L, int dtype);

/* StatBuffer::findMinMax -> extrema_scan (compressor.hpp:23-31,
 * compressor.cu:228-302).  d_minmax: double[4]; [0] = min and [1] = max on
 * return, [2..3] are scratch. */
int pa_extrema(const void* d_data, int dtype, int64_t n, double* d_minmax, void* stream);

/* Auto-tuning part of spline_construct (InterpolationPredictor.cu:2273-2355,
 * kernel auto_tuning_interpolation 880-967).  Scores the 15 candidates with a
 * fixed-order sum and picks flags on the device.  d_scores15 (double[15]) may
 * be NULL. */
size_t pa_autotune_workspace(const pa_layout* L);
int pa_autotune(const void* d_data, int dtype, const pa_layout* L, double* d_scores15,
                uint8_t* d_flags8, void* d_ws, size_t ws_bytes, void* stream);

/* pa_extrema followed by pa_autotune (StatBuffer::findMinMax,
 * compressor.hpp:23-31, and the auto-tuning part of spline_construct,
 * InterpolationPredictor.cu:2273-2355: the two reads of the field before the
 * prediction), run side by side in one launch, and with d_bounds non-NULL
 * pa_level_bounds(d_minmax, rel_eb, dtype, d_bounds) after them: the same
 * d_minmax, scores, flags and bounds as the separate calls.  d_data must be
 * 16-byte aligned; the workspace is pa_autotune's. */
int pa_extrema_autotune(const void* d_data, int dtype, const pa_layout* L, double* d_minmax, double rel_eb,
                        double* d_bounds, double* d_scores15, uint8_t* d_flags8, void* d_ws, size_t ws_bytes,
                        void* stream);

/* Prediction part of spline_construct (kernel interpolation,
 * InterpolationPredictor.cu:2137-2173): quantisation codes in level-major
 * order (d_qc: L->code_len int32, fully written) and the anchors (n4 T). */
int pa_spline_construct(const void* d_data, int dtype, const pa_layout* L, double eb,
                        double rel_eb, const uint8_t* d_flags8, int32_t* d_qc, void* d_anchors,
                        void* stream);

/* The host steps of the bound (eb = rel_eb * range, prism.cu:16-19; calc_eb,
 * InterpolationPredictor.cu:1825-1839) on the device, so that a compression
 * needs no host round trip before its last stage.  d_minmax: the result of
 * pa_extrema for a relative bound (eb relative), NULL for an absolute eb.
 * d_bounds (double[16]) = { absolute eb, range (1 for an absolute eb),
 * eb_r[4], ebx2[4] (float64), eb_r[4] | ebx2[4] (float32, packed in 4
 * doubles), status (0, or PA_ERR_ARG when the bound is not positive, e.g.
 * a constant field: the stages then run on a placeholder bound of 1), 0 }.
 * The values equal the host computation's. */
int pa_level_bounds(const double* d_minmax, double eb, int dtype, double* d_bounds, void* stream);
/* pa_spline_construct with the bounds of pa_level_bounds */
int pa_spline_construct_dev(const void* d_data, int dtype, const pa_layout* L, const double* d_bounds,
                            const uint8_t* d_flags8, int32_t* d_qc, void* d_anchors, void* stream);

/* spline_reconstruct (InterpolationPredictor.cu:2372-2421).  d_maxlevel is a
 * device int (levels with level_id > *d_maxlevel are skipped) or NULL for 4. */
int pa_spline_reconstruct(const int32_t* d_qc, const void* d_anchors, int dtype,
                          const pa_layout* L, double eb, double rel_eb, const uint8_t* d_flags8,
                          const int32_t* d_maxlevel, void* d_out, void* stream);

/* spline_progressive_reconstruct (InterpolationPredictor.cu:2423-2484):
 * out = old + reconstruction of delta codes with zero anchors. */
int pa_spline_delta_reconstruct(const int32_t* d_qc, int dtype, const pa_layout* L, double eb,
                                double rel_eb, const uint8_t* d_flags8,
                                const int32_t* d_maxlevel, const void* d_old, void* d_out,
                                void* stream);

/* convert_to_bitplane<E, NB|SM> (bitplane.cu:604-622).  d_bp: L->host_size
 * bytes, fully written (bytes between the device and host layouts are 0). */
int pa_bitshuffle(const int32_t* d_qc, const pa_layout* L, int mode, uint8_t* d_bp,
                  void* stream);
/* inverse_convert_to_bitplane (bitplane.cu:624-643) */
int pa_inverse_bitshuffle(const uint8_t* d_bp, const pa_layout* L, int mode, int32_t* d_qc,
                          void* stream);
/* inverse_convert_to_bitplane_progressive (bitplane.cu:645-667): only the
 * planes [begin[lr], end[lr]) of each level contribute (plus the sign plane
 * for PA_MODE_SM).  Only those planes are read. */
int pa_inverse_bitshuffle_window(const uint8_t* d_bp, const pa_layout* L, int mode,
                                 const int32_t* d_begin, const int32_t* d_end, int32_t* d_qc,
                                 void* stream);

/* Narrow hand-off between the inverse shuffle and the reconstruction of a
 * progressive round (no reference counterpart): when every level that reads
 * planes starts at plane 17 or later, the delta codes fit int16 and are
 * written as int16 at the same indices, halving the code traffic of both
 * stages; *d_narrow (device int32) records the choice and the _narrow
 * reconstructions read it.  Results equal the int32 path's. */
int pa_inverse_bitshuffle_window_narrow(const uint8_t* d_bp, const pa_layout* L, const int32_t* d_begin,
                                        const int32_t* d_end, void* d_qc, int32_t* d_narrow, void* stream);
int pa_spline_reconstruct_narrow(const void* d_qc, const int32_t* d_narrow, const void* d_anchors, int dtype,
                                 const pa_layout* L, double eb, double rel_eb, const uint8_t* d_flags8,
                                 const int32_t* d_maxlevel, void* d_out, void* stream);
int pa_spline_delta_reconstruct_narrow(const void* d_qc, const int32_t* d_narrow, int dtype, const pa_layout* L,
                                       double eb, double rel_eb, const uint8_t* d_flags8,
                                       const int32_t* d_maxlevel, const void* d_old, void* d_out, void* stream);

/* Progressive round without a code hand-off (no reference counterpart: the
 * inverse shuffle, bitplane.cu:645-667, folded into the reconstruction,
 * InterpolationPredictor.cu:2209-2239).  The deferred inverse shuffle acts
 * as the _narrow one, except that for a 3-D field (nz > 1) whose window
 * holds at most 8 planes on every level it writes no codes and sets *d_src
 * to 2; pa_spline_reconstruct_planes then reads each tile's codes from the
 * decoded planes d_bp (L->host_size bytes) under the window d_begin/d_end,
 * and otherwise from d_qc (int32 or int16 by *d_src, as the _narrow
 * reconstructions).  Exactly one of d_anchors (a field's first round) and
 * d_old (the delta form) is given.  Results equal the int32 path's. */
int pa_inverse_bitshuffle_window_deferred(const uint8_t* d_bp, const pa_layout* L, const int32_t* d_begin,
                                          const int32_t* d_end, void* d_qc, int32_t* d_src, void* stream);
int pa_spline_reconstruct_planes(const void* d_qc, const int32_t* d_src, const uint8_t* d_bp,
                                 const int32_t* d_begin, const int32_t* d_end, const void* d_anchors,
                                 const void* d_old, int dtype, const pa_layout* L, double eb, double rel_eb,
                                 const uint8_t* d_flags8, const int32_t* d_maxlevel, void* d_out, void* stream);

/* lossless_encode (zbpc.cu:476-520, kernel zbpc_encode 155-263): writes
 * header, size table and payloads to d_out (capacity >= pa_compress_bound),
 * the 4x32 cumulative size table to d_cum and the unpadded stream length to
 * d_outsize.  One pass over the chunks: each finds its payload offset by a
 * decoupled look-back over the chunks before it and writes its bytes
 * straight into the stream.  Workspace (8-byte aligned): an int32 error
 * word in an 8-byte slot (non-zero only if a chunk gave up waiting for the
 * sizes before it, which in-order dispatch rules out), then one 8-byte
 * status word per chunk. */
size_t pa_encode_workspace(const pa_layout* L);
int pa_lossless_encode(const uint8_t* d_bp, const pa_layout* L, uint8_t* d_out, uint64_t* d_cum,
                       uint64_t* d_outsize, void* d_ws, size_t ws_bytes, void* stream);

/* compress_merge (compressor.cu:139-151): pads the stream to 8 bytes and
 * appends flags, range, size table and anchors; d_total receives the final
 * stream length.  d_range is a device double (NULL: 1.0, absolute mode). */
int pa_compress_merge(uint8_t* d_out, const pa_layout* L, int dtype, const uint64_t* d_outsize,
                      const uint8_t* d_flags8, const double* d_range, const uint64_t* d_cum,
                      const void* d_anchors, uint64_t* d_total, void* stream);

/* lossless_decode (zbpc.cu:523-558, kernel zbpc_decode 265-365).  in_len is
 * the stream length in bytes; chunks whose table entries point outside it
 * are not read.  The first int32 of the workspace collects error bits
 * (non-zero: malformed stream; the affected chunks decode to zeros). */
size_t pa_decode_workspace(const pa_layout* L);
int pa_lossless_decode(const uint8_t* d_in, int64_t in_len, const pa_layout* L, uint8_t* d_bp,
                       void* d_ws, size_t ws_bytes, void* stream);
/* lossless_decode_progressive (zbpc.cu:560-599, kernel 368-473).  Decodes
 * every chunk that holds a byte of a requested device bit-plane (the covering
 * set).  The reference selects host chunks [begin*cb, end*cb) of each level;
 * because host planes are longer than device planes (io.hpp:104-132 vs
 * bitplane.cu:10-53) that selection misses the first bytes of a device plane
 * in rounds after the first.  d_bytes_read (may be NULL) accumulates the
 * payload bytes read.
 * The workspace must stay untouched between calls on the same stream when
 * offsets_ready != 0 (offsets are then reused). */
int pa_lossless_decode_window(const uint8_t* d_in, int64_t in_len, const pa_layout* L,
                              const int32_t* d_begin,
                              const int32_t* d_end, uint8_t* d_bp, uint64_t* d_bytes_read,
                              int offsets_ready, void* d_ws, size_t ws_bytes, void* stream);

/* Allocation-free pa_lossless_decode.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (error slot + chunk offsets, pa_decode_workspace(L) bytes) lives
 * in a fixed device array of pa_noalloc_scratch_bytes() = 1 MiB per GPU,
 * held until the process ends: fields of more than 131,064 chunks (host
 * bit-plane bytes above about 2 GiB) are rejected with PA_ERR_WORKSPACE
 * before any device work.  d_err (device int32) receives the decoder's
 * error word (0: well-formed).  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams or threads) race. */
size_t pa_noalloc_scratch_bytes(void);
int pa_lossless_decode_noalloc(const uint8_t* d_in, int64_t in_len, const pa_layout* L, uint8_t* d_bp,
                               int32_t* d_err, void* stream);

/* findStrategy_h (dataloader.cu:179-208): updates d_begin/d_end in place and
 * writes d_maxlevel.  eb and target_error are absolute.  One launch; the
 * workspace holds the search's best key and arrival count: zero-fill it
 * before its first use (every call leaves it at zero again), and do not
 * let calls that share a workspace overlap. */
size_t pa_strategy_workspace(void);
int pa_find_strategy(const uint64_t* d_cum, int32_t* d_begin, int32_t* d_end,
                     int32_t* d_maxlevel, double eb, double target_error, void* d_ws,
                     size_t ws_bytes, void* stream);

/* statistic (include/utils/analyze.hpp:5-70) on the device, in one pass:
 * d_out (double[9]) = { L-inf error, its first index, sum of squared errors,
 * min and max of d_ori, MSE, PSNR (dB), NRMSE, Max_RE }.  Deterministic
 * (fixed reduction order). */
size_t pa_quality_workspace(int64_t n);
int pa_quality(const void* d_ori, const void* d_rec, int dtype, int64_t n, double* d_out, void* d_ws,
               size_t ws_bytes, void* stream);

/* ---- Batched compression (no reference counterpart) ----
 * Many fields of mixed shapes and dtypes compressed by one fixed sequence of
 * launches (tuner, extrema, code gaps and encoder status clear in one; the
 * bounds and flags; spline per dtype; shuffle; encode; trailer: at most 7
 * kernels) whatever the number of fields, optionally captured once in a
 * HIP graph and replayed.
 * Each field's stream is byte-identical to the one the plain per-field
 * calls (pa_extrema .. pa_compress_merge) produce for the same input. */
typedef struct pa_batch_field {
    const void* data; /* device T[nz][ny][nx], 16-byte aligned */
    int32_t dtype;    /* PA_F32 / PA_F64 */
    int32_t mode;     /* PA_MODE_NB (progressive) / PA_MODE_SM */
    int32_t rel;      /* 1: eb is relative to the field's value range */
    int32_t reserved;
    double eb;
    uint32_t nx, ny, nz, reserved2;
    uint8_t* out; /* device, >= pa_compress_bound(layout, dtype) bytes */
} pa_batch_field;

typedef struct pa_batch_plan pa_batch_plan;

/* device workspace of a batch: per-field scratch plus the descriptor tables */
size_t pa_batch_workspace(const pa_batch_field* fields, int nf);
/* Plans a batch (host only; uploads the descriptor tables synchronously).
 * d_result: device uint64[4 * nf]; after each run, entry 4i.. of field i =
 * { stream length, status (0 or PA_ERR_ARG for a non-positive bound, e.g. a
 * constant field in relative mode), value range and absolute bound as f64
 * bit patterns }.  The tables are uploaded on `stream` (runs on another
 * stream must be ordered after it).  use_graph != 0 captures the launch
 * sequence once in a HIP graph that every run replays. */
int pa_batch_plan_create(const pa_batch_field* fields, int nf, void* d_ws, size_t ws_bytes,
                         uint64_t* d_result, int use_graph, void* stream, pa_batch_plan** plan);
/* Points the plan at new inputs / outputs of the same shapes, dtypes, modes
 * and bound kinds (bounds may change).  Synchronises `stream` first (earlier
 * runs read the tables being replaced), then uploads on it. */
int pa_batch_plan_rebind(pa_batch_plan* plan, const pa_batch_field* fields, int nf, void* stream);
int pa_batch_run(pa_batch_plan* plan, void* stream);
/* kernel launches one run issues (graph nodes when captured) */
int pa_batch_plan_launches(const pa_batch_plan* plan);
/* Launch-parameter autotuner: times every candidate of each tunable stage
 * on the plan's own fields (reps timed launches each; synchronises
 * `stream`), keeps the fastest and re-captures the graph.  Tuned: extrema
 * load batches per workgroup {2, 4, 8, 16}, spline work order (contiguous
 * per XCD or launch order), encoder chunks per workgroup {2, 4, 8}.  Streams are
 * unchanged by any choice. */
int pa_batch_plan_autotune(pa_batch_plan* plan, int reps, void* stream);
/* params3 = {extrema batches, XCD order, encoder chunks per workgroup};
 * timings9 (may be NULL) = the autotuner's ms per candidate in that order
 * (4 + 2 + 3) */
int pa_batch_plan_params(const pa_batch_plan* plan, int32_t* params3, float* timings9);
int pa_batch_plan_destroy(pa_batch_plan* plan);

/* ---- Batched progressive retrieval (no reference counterpart) ----
 * One round of decompress_progressive_pipeline (compressor.cu:166-190) for
 * each of many resident fields, by one fixed sequence of launches (strategy
 * search and update, offsets and window, count scan, decode, inverse
 * shuffle, reconstruction: at most 8 kernels) whatever the number of jobs.
 * Each job's state and output equal those of the plain calls
 * pa_find_strategy, pa_lossless_decode_window, pa_inverse_bitshuffle_window
 * and pa_spline_reconstruct (round 0) / pa_spline_delta_reconstruct. */
typedef struct pa_retrieve_job {
    const uint8_t* stream; /* device: the field's progressive stream */
    int64_t stream_len;
    int32_t dtype;
    uint32_t nx, ny, nz;
    double eb;        /* absolute bound of the stream */
    double rel_eb;    /* relative bound of the stream (0 in absolute mode) */
    double target;    /* absolute target error of this round */
    int32_t* state;   /* device int32[14], zeros before round 0: begin[4], end[4], maxlevel, code source
                         (0 int32, 1 int16, 2 bit-planes), each level's first live plane of the window */
    uint8_t* bp;      /* device, layout host_size bytes, kept between rounds */
    int32_t* qc;      /* device, code_len int32, kept between rounds */
    void* dws;        /* device, pa_decode_workspace bytes, kept between rounds */
    uint64_t* bytes_read; /* device, may be NULL: accumulates payload bytes read */
    int32_t round;    /* rounds already run on this field */
    const void* old;  /* device: the previous reconstruction (round > 0) */
    void* out;        /* device: the new reconstruction (n values) */
} pa_retrieve_job;

/* bytes of the job table staged from the host */
size_t pa_retrieve_table_bytes(int nj);
/* device workspace of a batch of nj jobs (job table, strategy keys, scan) */
size_t pa_retrieve_workspace(int nj);
/* h_table: pinned host memory of pa_retrieve_table_bytes(nj) bytes, left
 * untouched until the work on `stream` has completed (the table is copied
 * from it asynchronously); NULL: the call synchronises `stream` and copies
 * from memory of its own.  Jobs must name distinct fields. */
int pa_retrieve_batch(const pa_retrieve_job* jobs, int nj, void* h_table, size_t table_bytes, void* d_ws,
                      size_t ws_bytes, void* stream);

/* Library / device information (no reference counterpart). */
const char* pa_version(void);
int pa_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
