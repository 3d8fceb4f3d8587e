// analyze.hpp -- quality and throughput reports of the PRISM command line.
//
// Same figures and labels as the reference's include/utils/analyze.hpp:5-158:
// L-infinity error (Max_E) and its index, Max_E / range (Max_RE), PSNR =
// 20 log10(range) - 10 log10(MSE), NRMSE = sqrt(MSE) / range, per-stage times
// with GiB/s of original data, and, per progressive round, the payload bytes
// that the reference's accounting attributes to the planes retrieved.
//
// statistic, print_result and calculate_accumulate_size are at global scope
// with the reference's parameter lists; prism::calculate_accumulate_size adds
// the payload bytes the decoder actually read (the covering chunks).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>

#include <hip/hip_runtime.h>

template <typename T>
void statistic(const T* ori, const T* data, size_t n) {
    double mx = ori[0], mn = ori[0], max_err = 0.0, l2 = 0.0;
    size_t at = 0;
    for (size_t i = 0; i < n; ++i) {
        const double o = ori[i], e = std::fabs((double)data[i] - o);
        mx = o > mx ? o : mx;
        mn = o < mn ? o : mn;
        if (e > max_err) {
            max_err = e;
            at = i;
        }
        l2 += e * e;
    }
    const double range = mx - mn, mse = l2 / (double)n;
    std::printf("  Max_E  = %.15lf, idx = %d\n", max_err, (int)at);
    std::printf("  Max_RE = %G\n", max_err / range);
    std::printf("  PSNR = %.7lf, NRMSE = %.10G\n", 20 * std::log10(range) - 10 * std::log10(mse),
                std::sqrt(mse) / range);
}

// decmp = 0: compression stages; decmp = 1: decompression stages, plus the
// strategy search when progressive = 1.  Times in ms.  targeteb and
// targetreb are accepted for the reference's signature and not printed
// (analyze.hpp:75-77 does not print them either).
template <typename T, int decmp = 1, int progressive = 1>
void print_result(size_t size, size_t total_compressed_size, double targeteb, double targetreb, double time_enum,
                  double time_pred, double time_bitplane, double time_lossless) {
    (void)targeteb;
    (void)targetreb;
    std::printf("\n  \033[1m%-12s %-12s %-20s\033[0m\n", "type", "time (ms)", "throughtput (GB/s)");
    const double total_bytes = (double)size * sizeof(T);
    auto th = [total_bytes](double ms) { return total_bytes / 1024 / 1024 / 1024 / ms * 1000; };
    if (decmp == 1) {
        if (progressive == 1) std::printf("  %-12s %-12f %-10.4f\n", "loadStrategy", time_enum, th(time_enum));
        std::printf("  %-12s %-12f %-10.4f\n", "ipredict", time_pred, th(time_pred));
        std::printf("  %-12s %-12f %-10.4f\n", "ibitplane", time_bitplane, th(time_bitplane));
        std::printf("  %-12s %-12f %-10.4f\n", "ilossless", time_lossless, th(time_lossless));
        const double total = time_pred + time_bitplane + time_lossless + (progressive == 1 ? time_enum : 0.0);
        std::printf("  \033[1m%-12s\033[0m %-12f %-10.4f\n\n", "itotal", total, th(total));
    } else {
        std::printf("  %-12s %-12f %-10.4f\n", "predict", time_pred, th(time_pred));
        std::printf("  %-12s %-12f %-10.4f\n", "bitplane", time_bitplane, th(time_bitplane));
        std::printf("  %-12s %-12f %-10.4f\n", "lossless", time_lossless, th(time_lossless));
        const double total = time_pred + time_bitplane + time_lossless;
        std::printf("  \033[1m%-12s\033[0m %-12f %-10.4f\n\n", "total", total, th(total));
        std::printf("  compression ratio: %lf\n", total_bytes / (double)total_compressed_size);
        std::printf("--------------------------------------\n");
    }
}

namespace prism {

// Payload bytes of the planes [begin, end) of each level from the
// cumulative size table, the reference's accounting (analyze.hpp:118-158);
// loaded_size accumulates over rounds.  actually_read is the payload the
// decoder read this round (the covering chunks); a negative value omits
// that line.
template <typename T>
void calculate_accumulate_size(size_t& loaded_size, size_t total_size, const int* begin_d, const int* end_d,
                               const size_t* cum_d, long long actually_read) {
    int b0[4], b1[4];
    size_t cum[4][32];
    if (hipMemcpy(b0, begin_d, sizeof(b0), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b1, end_d, sizeof(b1), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(cum, cum_d, sizeof(cum), hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "[HIP ERROR] cannot read the retrieval state\n");
        return;
    }
    // cumulative -> per-plane sizes; plane 0 of a level follows the last
    // plane of the coarser level
    size_t per[4][32];
    for (int l = 0; l < 4; ++l)
        for (int b = 0; b < 32; ++b) {
            const size_t prev = b > 0 ? cum[l][b - 1] : (l > 0 ? cum[l - 1][31] : 0);
            per[l][b] = (l == 0 && b == 0) ? cum[0][0] : cum[l][b] - prev;
        }
    size_t round = 0;
    for (int l = 0; l < 4; ++l)
        for (int b = b0[l]; b < b1[l]; ++b) round += per[l][b];
    loaded_size += round;
    const double total_bytes = (double)total_size * sizeof(T);
    std::printf("  Data Chunk: %zu (%lf%%)\n", round, 100.0 * round / total_bytes);
    std::printf("  Total Retrieved size: %zu (%lf%%)\n", loaded_size, 100.0 * loaded_size / total_bytes);
    std::printf("  Bit rate: %lf\n", sizeof(T) * 8.0 * loaded_size / total_bytes);
    if (actually_read >= 0) std::printf("  Payload read by the decoder: %lld\n", actually_read);
    std::printf("--------------------------------------\n");
}

}  // namespace prism

// the reference's form (analyze.hpp:118)
template <typename T>
void calculate_accumulate_size(size_t& loaded_size, size_t total_size, int* begin_d, int* end_d,
                               size_t* compressedSize_bp_d) {
    prism::calculate_accumulate_size<T>(loaded_size, total_size, begin_d, end_d, compressedSize_bp_d, -1);
}
