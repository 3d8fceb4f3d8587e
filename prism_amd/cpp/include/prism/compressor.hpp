// compressor.hpp -- the PRISM C++ API (prism::extrema_scan, StatBuffer,
// Compressor<T,E>) over libprism_amd's C ABI.
//
// Names, public members and method signatures follow the reference's
// include/compressor.hpp:10-83; with io.hpp, parameters.hpp, prism.hpp,
// analyze.hpp, dataloader.hpp, timer.hpp and err.hpp beside it, a driver
// written against the reference's headers compiles and links against
// libprism_cpp.so once its CUDA runtime calls are spelled as their HIP
// counterparts (cudaStreamCreate -> hipStreamCreate, CHECK_CUDA ->
// CHECK_HIP): no CUDA name is provided here.  What differs:
//   - every stage runs on the caller's stream through one C-ABI call
//     (include/prism_amd.h); buffers and workspaces are allocated once in
//     init() and released in the destructor; nothing leaks;
//   - the interpolation flags, the size table, maxlevel and the stream
//     length live on the device; the host copies that the reference's API
//     exposes (intp_param, total_compressed_size, range) are refreshed at
//     the end of each pipeline call with one synchronisation;
//   - the progressive loader decodes the chunks that cover each requested
//     device bit-plane (see pa_lossless_decode_window).
// Errors from the HIP runtime or the library print a message and exit with
// EXIT_FAILURE, as the reference's CHECK_CUDA does (err.hpp:3-11).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>

#include "io.hpp"
#include "parameters.hpp"
#include "prism_amd.h"
#include "timer.hpp"

namespace prism {

// min and max of len device values of data into the host array result[2]
// (compressor.hpp:10-11, compressor.cu:228-302); one synchronisation.
template <typename T>
void extrema_scan(T* data, size_t len, T* result);

// A field plus its value range (compressor.hpp:17-31).
template <typename T>
struct StatBuffer : Buffer {
    double minVal{0.0}, maxVal{0.0}, range{1.0};
    StatBuffer() : Buffer() {}
    template <typename... Dim>
    StatBuffer(prism_dtype dtype, Dim... args) : Buffer(dtype, args...) {}
    // min, max and range of the device data (extrema kernel; one sync)
    double findMinMax();
};

template <typename T, typename E>
class Compressor {
   public:
    size_t len{0};
    size_t bytes_{0};
    int radius{0};
    int align{0};
    int HEADERSIZE{0};
    double range{1.0};
    // stage times of the last pipeline call, in ms (device events)
    double time_pred{0.0}, time_bitplane{0.0}, time_encode{0.0};
    double itime_pred{0.0}, itime_bitplane{0.0}, itime_decode{0.0}, itime_enum{0.0};
    inBuffer* qc{nullptr};               // quantisation codes, level-major
    inBuffer* ap{nullptr};               // anchors
    Buffer* compressed_data{nullptr};    // the stream
    Buffer* profiling_errors{nullptr};   // tuner scores (15 doubles)
    Bitplane* bp{nullptr};
    uint8_t* compressed_bp{nullptr};     // = compressed_data->d
    int *begin{nullptr}, *end{nullptr};  // the context's retrieval state
    size_t compressed_lossless_size{0};  // header + size table + payloads + pad
    size_t total_compressed_size{0};     // whole stream
    std::array<std::array<size_t, 32>, 4> compressedSize_bp{};
    size_t* compressedSize_bp_d{nullptr};  // cumulative payload size by (lr, plane)
    uint64_t bytes_read{0};                // payload bytes decoded by the last retrieval

    Compressor() {}
    Compressor(const Compressor&) = delete;
    Compressor& operator=(const Compressor&) = delete;
    ~Compressor();
    void init(prism_context* config);
    void compute_align();

    void compress_pipeline(context* config, StatBuffer<T>* input, void* stream);
    void compress_predict(context* config, StatBuffer<T>* input, void* stream);
    void compress_merge(context* config, double input_range, void* stream);

    void decompress_pipeline(context* config, StatBuffer<T>* output, void* stream);
    template <int init_ap>
    void decompress_scatter(context* config, void* stream);
    void decompress_predict(context* config, StatBuffer<T>* output, void* stream);

    void decompress_progressive_pipeline(context* config, StatBuffer<T>* output_old, StatBuffer<T>* output_new,
                                         double targetError, double lastError, void* stream);
    void decompress_progressive_predict(context* config, StatBuffer<T>* output_old, StatBuffer<T>* output_new,
                                        void* stream);

   private:
    pa_layout L{};
    uint8_t* d_flags{nullptr};      // use_md[4], reverse[4]
    double* d_minmax{nullptr};      // [min, max, range, scratch]
    uint64_t* d_sizes{nullptr};     // [lossless length, total length, bytes read]
    int32_t* d_maxlevel{nullptr};   // [maxlevel, code source of the round (0 int32, 1 int16, 2 the planes)]
    // one workspace allocation cut into the tuner's, the encoder's, the
    // decoder's (kept across retrieval rounds) and the strategy search's
    uint8_t* ws{nullptr};
    size_t ws_off[4]{}, ws_len[4]{};
    bool offsets_ready{false};
    hipEvent_t ev[6]{};
    void sync_flags_to_host(context* config, void* stream);
};

}  // namespace prism
