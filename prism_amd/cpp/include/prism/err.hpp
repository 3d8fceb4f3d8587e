// err.hpp -- runtime error check of the PRISM C++ API: CHECK_HIP(call)
// prints the failing call's location and message and exits with
// EXIT_FAILURE, as the reference's CHECK_CUDA does (include/utils/err.hpp:3-11)
// for the CUDA runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <iostream>

#define CHECK_HIP(call)                                                                                \
    do {                                                                                               \
        hipError_t err_ = (call);                                                                      \
        if (err_ != hipSuccess) {                                                                      \
            std::cerr << "[HIP ERROR] " << __FILE__ << ":" << __LINE__ << " - " << hipGetErrorString(err_) \
                      << std::endl;                                                                    \
            std::exit(EXIT_FAILURE);                                                                   \
        }                                                                                              \
    } while (0)
