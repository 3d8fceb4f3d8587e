// timer.hpp -- stage timing with HIP events (the reference's
// include/utils/timer.hpp:23-30: GPUTimer with start(stream) and stop(stream)
// returning milliseconds after waiting for the stop event).
#pragma once

#include <hip/hip_runtime.h>

struct GPUTimer {
    hipEvent_t beg{}, end{};
    GPUTimer() {
        (void)hipEventCreate(&beg);
        (void)hipEventCreate(&end);
    }
    ~GPUTimer() {
        (void)hipEventDestroy(beg);
        (void)hipEventDestroy(end);
    }
    GPUTimer(const GPUTimer&) = delete;
    GPUTimer& operator=(const GPUTimer&) = delete;
    void start(void* stream) { (void)hipEventRecord(beg, reinterpret_cast<hipStream_t>(stream)); }
    double stop(void* stream) {
        (void)hipEventRecord(end, reinterpret_cast<hipStream_t>(stream));
        (void)hipEventSynchronize(end);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, beg, end);
        return ms;
    }
};
