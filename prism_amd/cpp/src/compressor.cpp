// compressor.cpp -- prism::Buffer / StatBuffer / Bitplane / Compressor over
// the C ABI.  Stage order and stream contents follow the reference's
// src/compressor.cu:63-226 and src/utils/io.cu:10-127.
#include "prism/compressor.hpp"
#include "prism/dataloader.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <type_traits>

namespace prism {

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "[HIP ERROR] %s: %s\n", what, hipGetErrorString(e));
        std::exit(EXIT_FAILURE);
    }
}

void check_pa(int rc, const char* what) {
    if (rc != PA_OK) {
        std::fprintf(stderr, "[PRISM ERROR] %s failed with status %d\n", what, rc);
        std::exit(EXIT_FAILURE);
    }
}

static hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------------------ Buffer --
Buffer::~Buffer() {
    if (d) (void)hipFree(d);
    if (h) (void)hipHostFree(h);
}

void Buffer::allocate(size_t nbytes) {
    bytes = nbytes;
    cap = nbytes ? nbytes : 1;
    check_hip(hipHostMalloc(&h, cap), "hipHostMalloc");
    check_hip(hipMalloc(&d, cap), "hipMalloc");
    check_hip(hipMemset(d, 0, cap), "hipMemset");
}

void Buffer::reserve(size_t nbytes) {
    if (nbytes <= cap) return;
    if (d) check_hip(hipFree(d), "hipFree");
    if (h) check_hip(hipHostFree(h), "hipHostFree");
    d = h = nullptr;
    cap = nbytes;
    check_hip(hipHostMalloc(&h, cap), "hipHostMalloc");
    check_hip(hipMalloc(&d, cap), "hipMalloc");
}

template <itype file_type>
void Buffer::load_fromfile(const std::string& filename) {
    std::ifstream file(filename, std::ios::binary | std::ios::ate);
    if (!file) {
        std::cerr << "Failed to open file: " << filename << std::endl;
        std::exit(1);
    }
    const size_t length = (size_t)file.tellg();
    file.seekg(0, std::ios::beg);
    if (file_type == ori_File) {
        if (length < bytes) {
            std::cerr << "Error: Read " << bytes << " bytes, but expected " << length << " bytes." << std::endl;
            std::cerr << "File read incomplete or corrupted." << std::endl;
            std::exit(1);
        }
    } else {
        reserve(length);
        bytes = length;
    }
    file.read(reinterpret_cast<char*>(h), (std::streamsize)bytes);
}
template void Buffer::load_fromfile<ori_File>(const std::string&);
template void Buffer::load_fromfile<cmp_File>(const std::string&);

void Buffer::unload_tofile(const std::string& filename, typetofile tf) {
    if (tf == typetofile::deviceTofile) D2H();
    std::ofstream out(filename, std::ios::binary);
    if (!out) {
        std::cerr << "can't open file" << std::endl;
        return;
    }
    out.write(reinterpret_cast<const char*>(h), (std::streamsize)bytes);
    std::cout << "decompressed file written: " << filename << '\n';
}

void Buffer::unload_tofile(const std::string& filename, long long numBytes, typetofile tf) {
    if (tf == typetofile::deviceTofile) D2H(numBytes);
    std::ofstream out(filename, std::ios::binary);
    if (!out) {
        std::cerr << "can't open file" << std::endl;
        return;
    }
    std::cout << "compressed file written: " << filename << '\n';
    out.write(reinterpret_cast<const char*>(h), (std::streamsize)numBytes);
}

void Buffer::D2H(long long n) {
    check_hip(hipMemcpy(h, d, n < 0 ? bytes : (size_t)n, hipMemcpyDeviceToHost), "D2H");
}
void Buffer::H2D(long long n) {
    check_hip(hipMemcpy(d, h, n < 0 ? bytes : (size_t)n, hipMemcpyHostToDevice), "H2D");
}
void Buffer::D2H_async(void* st, long long n) {
    check_hip(hipMemcpyAsync(h, d, n < 0 ? bytes : (size_t)n, hipMemcpyDeviceToHost, S(st)), "D2H");
}
void Buffer::H2D_async(void* st, long long n) {
    check_hip(hipMemcpyAsync(d, h, n < 0 ? bytes : (size_t)n, hipMemcpyHostToDevice, S(st)), "H2D");
}

// -------------------------------------------------------------- StatBuffer --
template <typename T>
void extrema_scan(T* data, size_t len, T* result) {
    double* d_mm = nullptr;
    check_hip(hipMalloc((void**)&d_mm, 4 * sizeof(double)), "hipMalloc");
    check_pa(pa_extrema(data, sizeof(T) == 4 ? PA_F32 : PA_F64, (int64_t)len, d_mm, nullptr), "pa_extrema");
    double mm[2];
    check_hip(hipMemcpy(mm, d_mm, sizeof(mm), hipMemcpyDeviceToHost), "D2H");
    check_hip(hipFree(d_mm), "hipFree");
    result[0] = (T)mm[0];  // the extrema of T values are T values: exact
    result[1] = (T)mm[1];
}
template void extrema_scan<float>(float*, size_t, float*);
template void extrema_scan<double>(double*, size_t, double*);

template <typename T>
double StatBuffer<T>::findMinMax() {
    T result[2];
    extrema_scan<T>(reinterpret_cast<T*>(d), len, result);
    minVal = result[0];
    maxVal = result[1];
    range = maxVal - minVal;
    return range;
}
template struct StatBuffer<float>;
template struct StatBuffer<double>;

// ---------------------------------------------------------------- Bitplane --
Bitplane::Bitplane(uint32_t x, uint32_t y, uint32_t z) {
    check_pa(pa_layout_init(&L, x, y, z), "pa_layout_init");
    dtype = U1;
    tsize = 1;
    lx = x;
    ly = y;
    lz = z;
    len = (size_t)L.n;
    aligned_len = (size_t)(L.n + L.dev_align);
    aligned_size = (size_t)L.host_size;
    for (int l = 0; l < 4; ++l) {
        strides[l] = L.dev_stride[l];
        aligned_strides[l] = L.host_stride[l];
    }
    allocate(aligned_size);
}

// -------------------------------------------------------------- Compressor --
// Bytes of the code buffer beyond the n element codes: L.code_len = n + n4 +
// dev_align int32 (the anchor slots and the padding of the level regions
// come on top of the elements' codes).
static int code_buffer_extra(const pa_layout& L) { return (int)(4 * (L.n4 + L.dev_align)); }

}  // namespace prism

// Size in bytes of the code buffer Compressor::init allocates for a shape
// (0 for an invalid shape); host only, for checking it against the layout.
extern "C" size_t prism_code_buffer_bytes(uint32_t x, uint32_t y, uint32_t z) {
    pa_layout L;
    if (pa_layout_init(&L, x, y, z) != PA_OK) return 0;
    return 4 * (size_t)L.n + (size_t)prism::code_buffer_extra(L);
}

namespace prism {

template <typename T, typename E>
Compressor<T, E>::~Compressor() {
    delete qc;
    delete ap;
    delete compressed_data;
    delete profiling_errors;
    delete bp;
    for (void* p : {(void*)compressedSize_bp_d, (void*)d_flags, (void*)d_minmax, (void*)d_sizes,
                    (void*)d_maxlevel, (void*)ws})
        if (p) (void)hipFree(p);
    for (hipEvent_t& e : ev)
        if (e) (void)hipEventDestroy(e);
}

template <typename T, typename E>
void Compressor<T, E>::compute_align() {
    align = (int)L.dev_align;
}

template <typename T, typename E>
void Compressor<T, E>::init(prism_context* config) {
    static_assert(std::is_same<E, int32_t>::value, "codes are int32");
    const uint32_t x = config->x, y = config->y, z = config->z;
    check_pa(pa_layout_init(&L, x, y, z), "pa_layout_init");
    radius = (int)config->radius;
    len = (size_t)L.n;
    bytes_ = len * sizeof(T);
    compute_align();
    qc = new inBuffer(I4, code_buffer_extra(L), x, y, z);
    if (qc->bytes < 4 * (size_t)L.code_len) {
        std::fprintf(stderr, "[PRISM ERROR] code buffer %zu B < %lld B\n", qc->bytes, 4LL * L.code_len);
        std::exit(EXIT_FAILURE);
    }
    ap = new inBuffer(config->dtype, 0, (x + 15) / 16, (y + 15) / 16, (z + 15) / 16);
    compressed_data = new Buffer(U1, 0, (uint32_t)1);
    profiling_errors = new Buffer(F8, 0, 15);
    bp = new Bitplane(x, y, z);
    HEADERSIZE = (int)(L.trailer_fixed + L.n4 * (int64_t)sizeof(T));
    begin = config->begin;
    end = config->end;
    check_hip(hipMalloc((void**)&compressedSize_bp_d, 4 * 32 * sizeof(size_t)), "hipMalloc");
    check_hip(hipMalloc((void**)&d_flags, 8), "hipMalloc");
    check_hip(hipMalloc((void**)&d_minmax, 4 * sizeof(double)), "hipMalloc");
    check_hip(hipMalloc((void**)&d_sizes, 4 * sizeof(uint64_t)), "hipMalloc");
    check_hip(hipMalloc((void**)&d_maxlevel, 2 * sizeof(int32_t)), "hipMalloc");
    ws_len[0] = pa_autotune_workspace(&L);
    ws_len[1] = pa_encode_workspace(&L);
    ws_len[2] = pa_decode_workspace(&L);
    ws_len[3] = pa_strategy_workspace();
    size_t total = 0;
    for (int i = 0; i < 4; ++i) {
        ws_off[i] = total;
        total += (ws_len[i] + 255) / 256 * 256;
    }
    check_hip(hipMalloc((void**)&ws, total ? total : 256), "hipMalloc");
    // the strategy search's best key and arrival count start at zero (its calls leave them there)
    check_hip(hipMemset(ws + ws_off[3], 0, ws_len[3]), "hipMemset");
    for (hipEvent_t& e : ev) check_hip(hipEventCreate(&e), "hipEventCreate");
}

template <typename T, typename E>
void Compressor<T, E>::sync_flags_to_host(context* config, void* stream) {
    uint8_t f[8];
    check_hip(hipMemcpyAsync(f, d_flags, 8, hipMemcpyDeviceToHost, S(stream)), "D2H");
    check_hip(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
    for (int i = 0; i < 4; ++i) {
        config->intp_param.use_md[i] = f[i] != 0;
        config->intp_param.reverse[i] = f[4 + i] != 0;
    }
}

static double elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? (double)ms : 0.0;
}

template <typename T, typename E>
void Compressor<T, E>::compress_pipeline(context* config, StatBuffer<T>* input, void* stream) {
    compressed_data->reserve((size_t)pa_compress_bound(&L, sizeof(T)));
    check_hip(hipEventRecord(ev[0], S(stream)), "hipEventRecord");
    compress_predict(config, input, stream);
    check_hip(hipEventRecord(ev[1], S(stream)), "hipEventRecord");
    check_pa(pa_bitshuffle((const int32_t*)qc->d, &L, config->bt == NB ? PA_MODE_NB : PA_MODE_SM,
                           (uint8_t*)bp->d, stream),
             "pa_bitshuffle");
    check_hip(hipEventRecord(ev[2], S(stream)), "hipEventRecord");
    check_pa(pa_lossless_encode((const uint8_t*)bp->d, &L, (uint8_t*)compressed_data->d,
                                (uint64_t*)compressedSize_bp_d, d_sizes, ws + ws_off[1], ws_len[1], stream),
             "pa_lossless_encode");
    compress_merge(config, input->range, stream);
    check_hip(hipEventRecord(ev[3], S(stream)), "hipEventRecord");
    uint64_t sz[2];
    check_hip(hipMemcpyAsync(sz, d_sizes, sizeof(sz), hipMemcpyDeviceToHost, S(stream)), "D2H");
    check_hip(hipMemcpyAsync(compressedSize_bp.data(), compressedSize_bp_d, sizeof(compressedSize_bp),
                             hipMemcpyDeviceToHost, S(stream)),
              "D2H");
    int32_t enc_err = 0;
    check_hip(hipMemcpyAsync(&enc_err, ws + ws_off[1], sizeof(enc_err), hipMemcpyDeviceToHost, S(stream)), "D2H");
    sync_flags_to_host(config, stream);
    if (enc_err) {
        std::fprintf(stderr, "[PRISM ERROR] lossless_encode: a chunk gave up waiting for its payload offset (%d)\n",
                     enc_err);
        std::exit(EXIT_FAILURE);
    }
    compressed_lossless_size = (size_t)((sz[0] + 7) / 8 * 8);
    total_compressed_size = (size_t)sz[1];
    compressed_bp = (uint8_t*)compressed_data->d;
    time_pred = elapsed(ev[0], ev[1]);
    time_bitplane = elapsed(ev[1], ev[2]);
    time_encode = elapsed(ev[2], ev[3]);
}

template <typename T, typename E>
void Compressor<T, E>::compress_predict(context* config, StatBuffer<T>* input, void* stream) {
    const int dt = sizeof(T) == 4 ? PA_F32 : PA_F64;
    check_pa(pa_autotune(input->d, dt, &L, (double*)profiling_errors->d, d_flags, ws + ws_off[0], ws_len[0], stream),
             "pa_autotune");
    check_pa(pa_spline_construct(input->d, dt, &L, config->eb, config->rel_eb, d_flags, (int32_t*)qc->d, ap->d,
                                 stream),
             "pa_spline_construct");
}

template <typename T, typename E>
void Compressor<T, E>::compress_merge(context* config, double input_range, void* stream) {
    (void)config;
    // the range travels through the pinned mirror of profiling_errors (its
    // last slot), so the copy is a true asynchronous one; every pipeline
    // call ends synchronised, so no earlier copy still reads that slot
    double* hr = (double*)profiling_errors->h + 14;
    *hr = input_range;
    check_hip(hipMemcpyAsync(d_minmax + 2, hr, sizeof(double), hipMemcpyHostToDevice, S(stream)), "H2D");
    check_pa(pa_compress_merge((uint8_t*)compressed_data->d, &L, sizeof(T) == 4 ? PA_F32 : PA_F64, d_sizes, d_flags,
                               d_minmax + 2, (const uint64_t*)compressedSize_bp_d, ap->d, d_sizes + 1, stream),
             "pa_compress_merge");
}

template <typename T, typename E>
template <int init_ap>
void Compressor<T, E>::decompress_scatter(context* config, void* stream) {
    total_compressed_size = compressed_data->bytes;
    if (total_compressed_size < (size_t)HEADERSIZE + 8) {
        std::fprintf(stderr, "[PRISM ERROR] compressed stream too short (%zu bytes)\n", total_compressed_size);
        std::exit(EXIT_FAILURE);
    }
    compressed_lossless_size = total_compressed_size - HEADERSIZE;
    compressed_bp = (uint8_t*)compressed_data->d;
    const uint8_t* tr = compressed_bp + compressed_lossless_size;  // trailer (compressor.cu:139-151)
    check_hip(hipMemcpyAsync(d_flags, tr, 8, hipMemcpyDeviceToDevice, S(stream)), "D2D");
    check_hip(hipMemcpyAsync(compressedSize_bp_d, tr + 16, 4 * 32 * sizeof(size_t), hipMemcpyDeviceToDevice,
                             S(stream)),
              "D2D");
    if (init_ap == 1)
        check_hip(hipMemcpyAsync(ap->d, tr + 16 + 4 * 32 * 8, ap->bytes, hipMemcpyDeviceToDevice, S(stream)), "D2D");
    double* hr = (double*)profiling_errors->h + 14;
    check_hip(hipMemcpyAsync(hr, tr + 8, sizeof(double), hipMemcpyDeviceToHost, S(stream)), "D2H");
    check_hip(hipMemcpyAsync(compressedSize_bp.data(), compressedSize_bp_d, sizeof(compressedSize_bp),
                             hipMemcpyDeviceToHost, S(stream)),
              "D2H");
    sync_flags_to_host(config, stream);
    range = *hr;
}

template <typename T, typename E>
void Compressor<T, E>::decompress_pipeline(context* config, StatBuffer<T>* output, void* stream) {
    if (config->bt == NB)  // compressor.cu:155-162: this path always inverts sign-magnitude planes
        std::fprintf(stderr, "warning: a progressive (-prog) stream is decoded without -errors; "
                             "the sign-magnitude inverse is applied as in the reference\n");
    decompress_scatter<1>(config, stream);
    if (config->error_mode == REL) config->eb = config->rel_eb * range;
    check_hip(hipEventRecord(ev[0], S(stream)), "hipEventRecord");
    check_pa(pa_lossless_decode(compressed_bp, (int64_t)total_compressed_size, &L, (uint8_t*)bp->d,
                                ws + ws_off[2], ws_len[2], stream),
             "pa_lossless_decode");
    check_hip(hipEventRecord(ev[1], S(stream)), "hipEventRecord");
    check_pa(pa_inverse_bitshuffle((const uint8_t*)bp->d, &L, PA_MODE_SM, (int32_t*)qc->d, stream),
             "pa_inverse_bitshuffle");
    check_hip(hipEventRecord(ev[2], S(stream)), "hipEventRecord");
    decompress_predict(config, output, stream);
    check_hip(hipEventRecord(ev[3], S(stream)), "hipEventRecord");
    int32_t err = 0;
    check_hip(hipMemcpyAsync(&err, ws + ws_off[2], sizeof(err), hipMemcpyDeviceToHost, S(stream)), "D2H");
    check_hip(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
    if (err) {
        std::fprintf(stderr, "[PRISM ERROR] malformed compressed stream (decoder status %d)\n", err);
        std::exit(EXIT_FAILURE);
    }
    itime_decode = elapsed(ev[0], ev[1]);
    itime_bitplane = elapsed(ev[1], ev[2]);
    itime_pred = elapsed(ev[2], ev[3]);
}

template <typename T, typename E>
void Compressor<T, E>::decompress_predict(context* config, StatBuffer<T>* output, void* stream) {
    check_pa(pa_spline_reconstruct((const int32_t*)qc->d, ap->d, sizeof(T) == 4 ? PA_F32 : PA_F64, &L, config->eb,
                                   config->rel_eb, d_flags, nullptr, output->d, stream),
             "pa_spline_reconstruct");
}

template <typename T, typename E>
void Compressor<T, E>::decompress_progressive_pipeline(context* config, StatBuffer<T>* output_old,
                                                       StatBuffer<T>* output_new, double targetError,
                                                       double lastError, void* stream) {
    const bool first = lastError == 0;
    if (first) decompress_scatter<1>(config, stream);
    else decompress_scatter<0>(config, stream);
    if (config->error_mode == REL) {
        config->eb = config->rel_eb * range;
        targetError *= range;
    }
    check_hip(hipEventRecord(ev[0], S(stream)), "hipEventRecord");
    check_pa(pa_find_strategy((const uint64_t*)compressedSize_bp_d, begin, end, d_maxlevel, config->eb, targetError,
                              ws + ws_off[3], ws_len[3], stream),
             "pa_find_strategy");
    check_hip(hipEventRecord(ev[1], S(stream)), "hipEventRecord");
    check_hip(hipMemsetAsync(d_sizes + 2, 0, sizeof(uint64_t), S(stream)), "hipMemsetAsync");
    check_pa(pa_lossless_decode_window(compressed_bp, (int64_t)total_compressed_size, &L, begin, end,
                                       (uint8_t*)bp->d, d_sizes + 2, offsets_ready ? 1 : 0, ws + ws_off[2], ws_len[2],
                                       stream),
             "pa_lossless_decode_window");
    offsets_ready = true;
    check_hip(hipEventRecord(ev[2], S(stream)), "hipEventRecord");
    // the inverse shuffle writes int32 or int16 codes, or none when the
    // reconstruction reads this round's planes itself (d_maxlevel[1] says which)
    int32_t* d_src = d_maxlevel + 1;
    check_pa(pa_inverse_bitshuffle_window_deferred((const uint8_t*)bp->d, &L, begin, end, qc->d, d_src, stream),
             "pa_inverse_bitshuffle_window_deferred");
    check_hip(hipEventRecord(ev[3], S(stream)), "hipEventRecord");
    const int dt = sizeof(T) == 4 ? PA_F32 : PA_F64;
    const bool anchored = first || output_old == nullptr;
    check_pa(pa_spline_reconstruct_planes(qc->d, d_src, (const uint8_t*)bp->d, begin, end, anchored ? ap->d : nullptr,
                                          anchored ? nullptr : output_old->d, dt, &L, config->eb, config->rel_eb,
                                          d_flags, d_maxlevel, output_new->d, stream),
             "pa_spline_reconstruct_planes");
    check_hip(hipEventRecord(ev[4], S(stream)), "hipEventRecord");
    int32_t err = 0;
    check_hip(hipMemcpyAsync(&err, ws + ws_off[2], sizeof(err), hipMemcpyDeviceToHost, S(stream)), "D2H");
    check_hip(hipMemcpyAsync(&bytes_read, d_sizes + 2, sizeof(bytes_read), hipMemcpyDeviceToHost, S(stream)), "D2H");
    check_hip(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
    if (err) {
        std::fprintf(stderr, "[PRISM ERROR] malformed compressed stream (decoder status %d)\n", err);
        std::exit(EXIT_FAILURE);
    }
    itime_enum = elapsed(ev[0], ev[1]);
    itime_decode = elapsed(ev[1], ev[2]);
    itime_bitplane = elapsed(ev[2], ev[3]);
    itime_pred = elapsed(ev[3], ev[4]);
}

template <typename T, typename E>
void Compressor<T, E>::decompress_progressive_predict(context* config, StatBuffer<T>* output_old,
                                                      StatBuffer<T>* output_new, void* stream) {
    check_pa(pa_spline_delta_reconstruct((const int32_t*)qc->d, sizeof(T) == 4 ? PA_F32 : PA_F64, &L, config->eb,
                                         config->rel_eb, d_flags, d_maxlevel, output_old->d, output_new->d, stream),
             "pa_spline_delta_reconstruct");
}

template class Compressor<float, i4>;
template class Compressor<double, i4>;

}  // namespace prism

// ------------------------------------------------------------ dataloader --
template <typename E>
void findStrategy_h(size_t* compressedSize_bp_d, int* begin, int* end, double eb, double targetError, double& time,
                    void* stream) {
    using prism::check_hip;
    const size_t wsb = pa_strategy_workspace();
    void* ws = nullptr;
    int32_t* d_maxlevel = nullptr;
    check_hip(hipMalloc(&ws, wsb ? wsb : 8), "hipMalloc");
    check_hip(hipMalloc((void**)&d_maxlevel, sizeof(int32_t)), "hipMalloc");
    hipEvent_t a, b;
    check_hip(hipEventCreate(&a), "hipEventCreate");
    check_hip(hipEventCreate(&b), "hipEventCreate");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    check_hip(hipMemsetAsync(ws, 0, wsb, st), "hipMemsetAsync");  // the search's best key and count start at zero
    check_hip(hipEventRecord(a, st), "hipEventRecord");
    prism::check_pa(pa_find_strategy(reinterpret_cast<const uint64_t*>(compressedSize_bp_d), begin, end, d_maxlevel, eb,
                                     targetError, ws, wsb, stream),
                    "pa_find_strategy");
    check_hip(hipEventRecord(b, st), "hipEventRecord");
    check_hip(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0.f;
    check_hip(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    time = ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    check_hip(hipFree(ws), "hipFree");
    check_hip(hipFree(d_maxlevel), "hipFree");
}
template void findStrategy_h<int32_t>(size_t*, int*, int*, double, double, double&, void*);
