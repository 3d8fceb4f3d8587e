// dataloader.hpp -- the loader's strategy search as a free function, with
// the reference's signature (include/loader/dataloader.hpp:9; dataloader.cu:
// 179-208): picks the planes [begin, end) of each level to read next so
// that the reconstruction error stays within targetError, updating the
// device arrays begin[4] / end[4] in place; time receives the device time
// in ms.  Runs on `stream` and waits for it.  The reference keeps the chosen
// maximum level in a process-global device variable; this build returns it
// through the Compressor (decompress_progressive_pipeline), so the free
// function has no such global.
#pragma once

#include <cstddef>

template <typename E>
void findStrategy_h(size_t* compressedSize_bp_d, int* begin, int* end, double eb, double targetError, double& time,
                    void* stream);
