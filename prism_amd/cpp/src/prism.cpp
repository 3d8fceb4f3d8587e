// prism.cpp -- the PRISM driver: compress a raw field file into a stream
// file, decompress it, or retrieve it progressively at a list of bounds
// (the flows of the reference's src/prism.cu:9-190).
//
// One stream and one Compressor serve all rounds of a progressive run: the
// compressed file is read once and the decoded bit-planes, codes and
// previous reconstruction stay resident in HBM between rounds (the
// reference reloads the file and rebuilds its Compressor each round,
// prism.cu:131-135, with the same results).
#include "prism/prism.hpp"

#include <cstdio>
#include <memory>

#include "prism/analyze.hpp"
#include "prism/compressor.hpp"

namespace {

using prism::check_hip;

template <typename T>
std::unique_ptr<prism::StatBuffer<T>> field(const prism_context* c) {
    return std::make_unique<prism::StatBuffer<T>>(c->dtype, 0, c->x, c->y, c->z);
}

template <typename T>
void prism_compress(prism_context* c, void* stream) {
    auto input = field<T>(c);
    input->template load_fromfile<ori_File>(c->oriFilePath);
    input->H2D();
    if (c->error_mode == REL) c->eb = c->rel_eb * input->findMinMax();
    if (c->report_time) {  // warm-up pass
        prism::Compressor<T, i4> warm;
        warm.init(c);
        warm.compress_pipeline(c, input.get(), stream);
    }
    prism::Compressor<T, i4> cmp;
    cmp.init(c);
    cmp.compress_pipeline(c, input.get(), stream);
    cmp.compressed_data->unload_tofile(c->cmpFilePath, (long long)cmp.total_compressed_size,
                                       prism::typetofile::deviceTofile);
    if (c->report_time)
        print_result<T, 0, 0>(c->size, cmp.total_compressed_size, c->eb, c->rel_eb, cmp.itime_enum, cmp.time_pred,
                              cmp.time_bitplane, cmp.time_encode);
}

template <typename T>
void load_stream(prism::Compressor<T, i4>& d, const prism_context* c) {
    d.compressed_data->template load_fromfile<cmp_File>(c->cmpFilePath);
    d.compressed_data->H2D();
}

// statistic (analyze.hpp:5-70) on the device: the original is uploaded once
// and compared with the reconstruction where it lies, in HBM
template <typename T>
void report_quality(const prism_context* c, prism::StatBuffer<T>* out) {
    auto ori = field<T>(c);
    ori->template load_fromfile<ori_File>(c->oriFilePath);
    ori->H2D();
    const int64_t n = (int64_t)c->size;
    const size_t wsb = pa_quality_workspace(n);
    void* ws = nullptr;
    double* d_q = nullptr;
    check_hip(hipMalloc(&ws, wsb), "hipMalloc");
    check_hip(hipMalloc((void**)&d_q, 9 * sizeof(double)), "hipMalloc");
    prism::check_pa(pa_quality(ori->d, out->d, sizeof(T) == 4 ? PA_F32 : PA_F64, n, d_q, ws, wsb, nullptr), "pa_quality");
    double q[9];
    check_hip(hipMemcpy(q, d_q, sizeof(q), hipMemcpyDeviceToHost), "D2H");
    check_hip(hipFree(ws), "hipFree");
    check_hip(hipFree(d_q), "hipFree");
    std::printf("  Max_E  = %.15lf, idx = %d\n", q[0], (int)q[1]);
    std::printf("  Max_RE = %G\n", q[8]);
    std::printf("  PSNR = %.7lf, NRMSE = %.10G\n", q[6], q[7]);
}

template <typename T>
void prism_decompress(prism_context* c, void* stream) {
    if (c->report_time) {  // warm-up pass
        auto tmp = field<T>(c);
        prism::Compressor<T, i4> warm;
        warm.init(c);
        load_stream(warm, c);
        warm.decompress_pipeline(c, tmp.get(), stream);
    }
    prism::Compressor<T, i4> d;
    d.init(c);
    load_stream(d, c);
    auto out = field<T>(c);
    d.decompress_pipeline(c, out.get(), stream);
    out->D2H();
    out->unload_tofile(c->decFilePath, prism::typetofile::hostTofile);
    if (c->report_time)
        print_result<T, 1, 0>(c->size, d.total_compressed_size, c->eb, c->rel_eb, d.itime_enum, d.itime_pred,
                              d.itime_bitplane, d.itime_decode);
    if (c->report_cr) report_quality<T>(c, out.get());
}

template <typename T>
void prism_progressive_decompress(prism_context* c, void* stream) {
    if (c->report_time) {  // warm-up round, then the retrieval state is reset
        auto tmp = field<T>(c);
        prism::Compressor<T, i4> warm;
        warm.init(c);
        load_stream(warm, c);
        warm.decompress_progressive_pipeline(c, nullptr, tmp.get(), c->target_ebs[0], 0, stream);
        check_hip(hipMemset(c->begin, 0, 4 * sizeof(int)), "hipMemset");
        check_hip(hipMemset(c->end, 0, 4 * sizeof(int)), "hipMemset");
    }
    prism::Compressor<T, i4> d;
    d.init(c);
    load_stream(d, c);
    auto cur = field<T>(c), prev = field<T>(c);
    size_t loaded = d.ap->bytes;
    const size_t rounds = c->target_ebs.size();
    for (size_t i = 0; i < rounds; ++i) {
        std::swap(cur, prev);  // prev holds the last reconstruction
        d.decompress_progressive_pipeline(c, i ? prev.get() : nullptr, cur.get(), c->target_ebs[i],
                                          i ? c->target_ebs[i - 1] : 0.0, stream);
        cur->D2H();
        if (c->report_time)
            print_result<T, 1, 1>(c->size, d.total_compressed_size, c->eb, c->rel_eb, d.itime_enum, d.itime_pred,
                                  d.itime_bitplane, d.itime_decode);
        if (c->report_cr) {
            report_quality<T>(c, cur.get());
            prism::calculate_accumulate_size<T>(loaded, c->size, c->begin, c->end, d.compressedSize_bp_d,
                                                (long long)d.bytes_read);
        }
    }
    cur->unload_tofile(c->decFilePath, prism::typetofile::hostTofile);
}

template <typename T>
void run(prism_context* c, void* stream) {
    if (c->isComp) prism_compress<T>(c, stream);
    if (c->isDecomp) {
        if (c->compMode && !c->target_ebs.empty()) prism_progressive_decompress<T>(c, stream);
        else prism_decompress<T>(c, stream);
    }
}

}  // namespace

void initialize(prism_context* config) {
    hipStream_t stream;
    check_hip(hipStreamCreate(&stream), "hipStreamCreate");
    if (config->dtype == PRISM_TYPE_FLOAT) run<f4>(config, stream);
    else if (config->dtype == PRISM_TYPE_DOUBLE) run<f8>(config, stream);
    check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    check_hip(hipStreamDestroy(stream), "hipStreamDestroy");
}
