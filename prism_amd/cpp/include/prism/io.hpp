// io.hpp -- device buffers of the PRISM C++ API: prism::Buffer (inBuffer)
// and prism::Bitplane, with the members and methods of the reference's
// include/utils/io.hpp:17-149 (file load / store, host <-> device copies,
// the bit-plane strides).  Each Buffer owns one device allocation and a
// pinned host mirror of the same size.
#pragma once

#include <hip/hip_runtime.h>

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

}  // namespace prism
