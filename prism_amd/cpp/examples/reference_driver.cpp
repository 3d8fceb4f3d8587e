// reference_driver.cpp -- a driver written against the PRISM C++ API names
// only, with the structure of the reference's src/prism.cu:9-190 and
// examples/prism.cpp: parse_argv, Compressor<T,i4> with a new compressor,
// stream reload and output buffer per progressive round, extrema_scan /
// findMinMax, print_result, statistic and calculate_accumulate_size on host
// arrays, findStrategy_h and GPUTimer.  It is built against the headers in
// prism_amd/cpp/include/prism (added to the include path, as the reference
// adds its include/ directories) and linked with libprism_cpp.so; the only
// spelling that differs from a CUDA driver is the runtime's (hip* calls,
// CHECK_HIP).  Used by the tests to show that such a caller runs unchanged
// and writes the same files as the oracle.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "analyze.hpp"
#include "compressor.hpp"
#include "dataloader.hpp"
#include "err.hpp"
#include "io.hpp"
#include "prism.hpp"
#include "timer.hpp"

template <typename T>
void compress(prism_context* config, void* stream) {
    auto* cmp = new prism::Compressor<T, i4>();
    auto input = new prism::StatBuffer<T>(config->dtype, 0, config->x, config->y, config->z);
    input->template load_fromfile<ori_File>(config->oriFilePath);
    input->H2D();
    T mm[2];
    prism::extrema_scan<T>(reinterpret_cast<T*>(input->d), input->len, mm);
    if (config->error_mode == REL) {
        config->eb = config->rel_eb * input->findMinMax();
        if (input->minVal != mm[0] || input->maxVal != mm[1]) std::printf("extrema disagree\n");
    }
    GPUTimer timer;
    timer.start(stream);
    cmp->init(config);
    cmp->compress_pipeline(config, input, stream);
    const double ms = timer.stop(stream);
    cmp->compressed_data->unload_tofile(config->cmpFilePath, cmp->total_compressed_size,
                                        prism::typetofile::deviceTofile);
    if (config->report_time == 1)
        print_result<T, 0, 0>(config->size, cmp->total_compressed_size, config->eb, config->rel_eb, cmp->itime_enum,
                              cmp->time_pred, cmp->time_bitplane, cmp->time_encode);
    std::printf("compress pipeline: %.3f ms\n", ms);
    delete cmp;
    delete input;
}

template <typename T>
void progressive(prism_context* config, void* stream) {
    prism::StatBuffer<T>* input = nullptr;
    if (config->report_cr == 1) {
        input = new prism::StatBuffer<T>(config->dtype, 0, config->x, config->y, config->z);
        input->template load_fromfile<ori_File>(config->oriFilePath);
    }
    prism::StatBuffer<T>* output_old = nullptr;
    prism::StatBuffer<T>* output_new = nullptr;
    size_t loaded_size = 0;
    for (size_t i = 0; i < config->target_ebs.size(); ++i) {
        delete output_old;
        output_old = output_new;
        output_new = new prism::StatBuffer<T>(config->dtype, 0, config->x, config->y, config->z);
        auto* decmp = new prism::Compressor<T, i4>();
        decmp->init(config);
        decmp->compressed_data->template load_fromfile<cmp_File>(config->cmpFilePath);
        decmp->compressed_data->H2D();
        if (i == 0) loaded_size = decmp->ap->bytes;
        decmp->decompress_progressive_pipeline(config, output_old, output_new, config->target_ebs[i],
                                               i ? config->target_ebs[i - 1] : 0, stream);
        output_new->D2H();
        if (config->report_time == 1)
            print_result<T, 1, 1>(config->size, decmp->total_compressed_size, config->eb, config->rel_eb,
                                  decmp->itime_enum, decmp->itime_pred, decmp->itime_bitplane, decmp->itime_decode);
        if (config->report_cr == 1) {
            statistic<T>(static_cast<T*>(input->h), static_cast<T*>(output_new->h), config->size);
            calculate_accumulate_size<T>(loaded_size, config->size, config->begin, config->end,
                                         decmp->compressedSize_bp_d);
        }
        if (i + 1 == config->target_ebs.size()) {
            // the strategy search as a free function, on a copy of the state:
            // the planes already read satisfy the last target, so nothing moves
            int *b = nullptr, *e = nullptr;
            CHECK_HIP(hipMalloc((void**)&b, 4 * sizeof(int)));
            CHECK_HIP(hipMalloc((void**)&e, 4 * sizeof(int)));
            CHECK_HIP(hipMemcpy(b, config->begin, 4 * sizeof(int), hipMemcpyDeviceToDevice));
            CHECK_HIP(hipMemcpy(e, config->end, 4 * sizeof(int), hipMemcpyDeviceToDevice));
            double t = 0;
            const double target = config->target_ebs[i] * (config->error_mode == REL ? decmp->range : 1.0);
            findStrategy_h<i4>(decmp->compressedSize_bp_d, b, e, config->eb, target, t, stream);
            int hb[4], he[4], ob[4], oe[4];
            CHECK_HIP(hipMemcpy(hb, b, sizeof(hb), hipMemcpyDeviceToHost));
            CHECK_HIP(hipMemcpy(he, e, sizeof(he), hipMemcpyDeviceToHost));
            CHECK_HIP(hipMemcpy(ob, config->end, sizeof(ob), hipMemcpyDeviceToHost));
            for (int l = 0; l < 4; ++l) oe[l] = ob[l];
            bool same = true;
            for (int l = 0; l < 4; ++l) same = same && hb[l] == oe[l] && he[l] == oe[l];
            std::printf("findStrategy_h after the last round: %s\n", same ? "no new planes" : "new planes");
            CHECK_HIP(hipFree(b));
            CHECK_HIP(hipFree(e));
            output_new->unload_tofile(config->decFilePath, prism::typetofile::hostTofile);
        }
        delete decmp;
    }
    delete output_old;
    delete output_new;
    delete input;
}

int main(int argc, char* argv[]) {
    prism_context* config = new prism_context;
    parse_argv(config, argc, argv);
    hipStream_t stream;
    CHECK_HIP(hipStreamCreate(&stream));
    if (config->dtype == PRISM_TYPE_FLOAT) {
        if (config->isComp) compress<f4>(config, stream);
        if (config->isDecomp) progressive<f4>(config, stream);
    } else {
        if (config->isComp) compress<f8>(config, stream);
        if (config->isDecomp) progressive<f8>(config, stream);
    }
    CHECK_HIP(hipStreamDestroy(stream));
    release_context(config);
    delete config;
    return 0;
}
