// compressor.hpp -- the PRISM C++ API (prism::Buffer, StatBuffer, Bitplane,
// Compressor<T,E>) over libprism_amd's C ABI.
//
// Class names, public members and method names follow the reference's
// include/compressor.hpp:35-90 and include/utils/io.hpp:17-149, so a driver
// written for it (src/prism.cu) compiles against this header.  What differs:
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

#include "parameters.hpp"
#include "prism_amd.h"

namespace prism {

// exits with a message when a HIP or library call fails
void check_hip(hipError_t e, const char* what);
void check_pa(int rc, const char* what);

enum class typetofile { hostTofile, deviceTofile };

// A device allocation of len elements of dtype plus `align` bytes, with a
// pinned host mirror (io.hpp:17-71).  Dimensions are x fastest.
typedef struct Buffer {
    prism_dtype dtype{U1};
    int tsize{1}, ndim{-1};
    size_t len{1}, bytes{1};
    uint32_t lx{1}, ly{1}, lz{1};
    size_t sty{1}, stz{1};
    int align{0};
    void* d{nullptr};
    void* h{nullptr};

    Buffer() {}
    template <typename... Dim>
    Buffer(prism_dtype type, int align_, Dim... args) : dtype(type), align(align_) {
        const uint32_t dims[] = {static_cast<uint32_t>(args)...};
        constexpr size_t nd = sizeof...(Dim);
        if (nd >= 1) lx = dims[0];
        if (nd >= 2) ly = dims[1 < nd ? 1 : 0];
        if (nd >= 3) lz = dims[2 < nd ? 2 : 0];
        ndim = lz > 1 ? 3 : (ly > 1 ? 2 : 1);
        tsize = dtype % 10;
        len = (size_t)lx * ly * lz;
        sty = lx;
        stz = (size_t)lx * ly;
        allocate((size_t)tsize * len + align);
    }
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    virtual ~Buffer();

    // raw file into the pinned host mirror: an original-data file must hold
    // at least `bytes` (the first `bytes` are read); a compressed file sets
    // `bytes` to the file length (io.cu:18-45)
    template <itype file_type>
    void load_fromfile(const std::string& filename);
    void unload_tofile(const std::string& filename, typetofile tf = typetofile::hostTofile);
    void unload_tofile(const std::string& filename, long long numBytes, typetofile tf = typetofile::hostTofile);
    void H2D(long long numBytes = -1);
    void D2H(long long numBytes = -1);
    void H2D_async(void* stream, long long numBytes = -1);
    void D2H_async(void* stream, long long numBytes = -1);
    // the reference's names for the two calls above
    void H2D_cudaasync(void* stream, long long numBytes = -1) { H2D_async(stream, numBytes); }
    void D2H_cudaasync(void* stream, long long numBytes = -1) { D2H_async(stream, numBytes); }

    template <typename D>
    D len3() const {
        return D(lx, ly, lz);
    }
    template <typename D>
    D st3() const {
        return D(1, sty, stz);
    }
    // grow both sides to at least nbytes (contents are not kept)
    void reserve(size_t nbytes);
    size_t capacity() const { return cap; }

   protected:
    size_t cap{0};  // bytes allocated on both sides
    void allocate(size_t nbytes);
} inBuffer;

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

// The bit-plane buffer of a field (io.hpp:75-149, io.cu:108-127): strides[]
// are the bytes per plane of the shuffle's layout, aligned_strides[] those of
// the layout that cuts the codec chunks; both indexed by level_id, 0 = unit 1.
// aligned_size = 32 * sum(aligned_strides) is the buffer size.
struct Bitplane : Buffer {
    pa_layout L{};
    size_t aligned_len{0}, aligned_size{0};
    int64_t strides[4]{0, 0, 0, 0};
    int64_t aligned_strides[4]{0, 0, 0, 0};
    Bitplane(uint32_t x, uint32_t y, uint32_t z);
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
    int32_t* d_maxlevel{nullptr};
    // one workspace allocation cut into the tuner's, the encoder's, the
    // decoder's (kept across retrieval rounds) and the strategy search's
    uint8_t* ws{nullptr};
    size_t ws_off[4]{}, ws_len[4]{};
    bool offsets_ready{false};
    hipEvent_t ev[6]{};
    void sync_flags_to_host(context* config, void* stream);
};

}  // namespace prism
