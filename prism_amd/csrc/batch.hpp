// batch.hpp -- device tables of a batch of fields and the stage launchers
// that the batched compression (batch.cpp) calls.  Every launcher issues a
// fixed number of kernels whatever the number of fields; a workgroup finds
// its field through the stage's prefix of per-field work counts.
#pragma once

#include "common.hpp"

namespace pa {

// Fields [0, nf32) hold float32 data and [nf32, nf) float64, so each
// dtype-specialised kernel covers one contiguous range of every prefix.
struct BatchDev {
    BatchField* fields;  // device array of nf descriptors
    int nf, nf32;
    const int64_t* p_ext;    // extrema workgroups
    const int64_t* p_tune;   // tuner sample blocks (x 15 candidates in grid.y)
    const int64_t* p_gap;    // code-gap workgroups
    const int64_t* p_tile;   // spline work units (a tile; 4 tiles of a 2-D, 256 of a 1-D field)
    const int64_t* p_shuf;   // shuffle workgroups
    const int64_t* p_chunk;  // encoder launch slots (one per wave64): chunks rounded up to a multiple of 8
    const int64_t* p_merge;  // trailer workgroups
    // host copies of the totals, and of the f32 part where a kernel is
    // specialised by dtype
    int64_t n_ext, n_tune, n_gap, n_tile, n_shuf, n_chunk, n_merge;
    int64_t n_tune32, n_tile32;
    // launch parameters (the plan's autotuner may change them)
    int ext_batches;  // 2048-vector load batches per extrema workgroup (p_ext is built for it)
    int encode_waves; // chunks (waves) per encoder workgroup
    int xcd_order;    // 1: spline work units dealt to XCDs in contiguous ranges; 0: launch order
};

constexpr int kExtChoices[4] = {2, 4, 8, 16};  // ext_batches candidates
constexpr int kEncodeChoices[3] = {2, 4, 8};   // encode_waves candidates

// per-field work counts of each stage (host)
int64_t batch_ext_blocks(const pa_layout* L, int dtype, int ext_batches);
int64_t batch_gap_blocks(const BatchField& f);
int64_t batch_shuf_blocks(const pa_layout* L);
int64_t batch_tile_units(const pa_layout* L);  // spline workgroups
int64_t batch_merge_blocks(const pa_layout* L, int dtype);
// host-side preparation of a descriptor (layout filled): tuner sampling,
// alpha, code gaps
void batch_predict_setup(BatchField& f);

// A batch of retrieval rounds: jobs [0, nj32) hold float32 fields.
struct RetrDev {
    const RetrJob* jobs;
    int nj, nj32;
    const int64_t* p_unshuf;  // inverse-shuffle workgroups
    const int64_t* p_unit;    // reconstruction work units
    int64_t* p_chunk;         // selected chunks (device-computed exclusive prefix, nj + 1)
    int64_t n_unshuf, n_unit, n_unit32, max_chunks;
};

int retr_strategy(const RetrDev& r, hipStream_t st);     // strategy.hip
int retr_decode(const RetrDev& r, hipStream_t st);       // codec.hip: offsets, window, decode
int retr_unshuffle(const RetrDev& r, hipStream_t st);    // bitplane.hip
int retr_reconstruct(const RetrDev& r, hipStream_t st);  // predictor.hip
// host helpers for job set-up
void retr_level_ebs(RetrJob& j, double eb, double rel_eb);
int64_t retr_unshuf_blocks(const pa_layout* L);

// The batched compression: batch_front, batch_spline, batch_shuffle,
// batch_encode (at most 7 launches).
int batch_front(const BatchDev& b, hipStream_t st);    // predictor.hip: tuner, extrema, code gaps and encoder
                                                       // status clear in one launch, then bounds and flags
int batch_spline(const BatchDev& b, hipStream_t st);   // predictor.hip: one launch per dtype
int batch_shuffle(const BatchDev& b, hipStream_t st);  // bitplane.hip
int batch_encode(const BatchDev& b, hipStream_t st);   // codec.hip: encode, trailer
int batch_keys_init(const BatchDev& b, hipStream_t st);  // predictor.hip: extrema keys before the first run
// the encoder alone, status words cleared first (the autotuner times it
// apart from the front launch that clears them in a run)
int batch_encode_chunks(const BatchDev& b, hipStream_t st);

}  // namespace pa
