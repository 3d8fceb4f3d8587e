// layout.cpp -- stream geometry of one field (host side).
//
// Reproduces the sizes of the reference: level grids and the device code
// layout with 8-code aligned level regions (InterpolationPredictor.cu:969-995,
// bitplane.cu:10-53), and the host bit-plane strides that cut codec chunks,
// which always pad by 1..8 codes (io.hpp:104-132, io.cu:292-302).
#include <string.h>

#include "prism_amd.h"

namespace {

int64_t device_plane_bytes(int64_t codes) {
    // (((codes)/8 - 1)/8 + 1)*8 with C division: an empty segment keeps 8 bytes
    int64_t groups = codes / 8;
    return groups == 0 ? 8 : ((groups - 1) / 8 + 1) * 8;
}

}  // namespace

extern "C" int pa_layout_init(pa_layout* L, uint32_t nx, uint32_t ny, uint32_t nz) {
    if (!L || nx == 0 || ny == 0 || nz == 0) return PA_ERR_ARG;
    const int64_t n = (int64_t)nx * ny * nz;
    if (n >= ((int64_t)1 << 31)) return PA_ERR_ARG;  // int32 element indices in the stream format
    memset(L, 0, sizeof(*L));
    L->nx = nx;
    L->ny = ny;
    L->nz = nz;
    L->n = n;

    int64_t d[3] = {nx, ny, nz};
    for (int lv = 0; lv <= 4; ++lv) {
        L->lx[lv] = d[0];
        L->lxy[lv] = d[0] * d[1];
        if (lv < 4) {
            for (int a = 0; a < 3; ++a) d[a] = (d[a] + 1) / 2;
            L->cnt[lv] = d[0] * d[1] * d[2];
        }
    }
    L->n4 = L->cnt[3];

    // level regions (absolute, anchors first): each level's count rounded up to 8
    const int64_t n4 = L->n4;
    int64_t count[4] = {n - L->cnt[0], L->cnt[0] - L->cnt[1], L->cnt[1] - L->cnt[2], L->cnt[2] - n4};
    L->prefix[4] = 0;
    L->prefix[3] = n4;
    int64_t pos = n4;
    for (int lv = 3; lv >= 1; --lv) {
        int64_t padded = (count[lv] + 7) / 8 * 8;
        L->dev_len[lv] = padded;
        pos += padded;
        L->prefix[lv - 1] = pos;
    }
    L->real_end = pos + count[0];
    // level 0 carries the n4 phantom codes of the shuffle (bitplane.cu:35)
    L->dev_len[0] = (count[0] + n4 + 7) / 8 * 8;
    L->dev_align = (L->prefix[0] - L->cnt[0]) + (L->dev_len[0] - count[0] - n4);
    L->code_len = n4 + n + L->dev_align;
    // the kernels index the code array with 32-bit arithmetic
    if (L->code_len >= ((int64_t)1 << 31)) return PA_ERR_ARG;

    for (int lv = 0; lv < 4; ++lv) L->dev_stride[lv] = device_plane_bytes(L->dev_len[lv]);
    int64_t base = 0;
    for (int lv = 3; lv >= 0; --lv) {
        L->dev_base[lv] = base;
        base += 32 * L->dev_stride[lv];
    }
    L->dev_size = base;

    // host strides: 1..8 codes of padding per level, then 8-byte rounding
    base = 0;
    for (int lv = 3; lv >= 0; --lv) {
        int64_t c = lv == 0 ? count[0] + n4 : count[lv];
        int64_t padded = c + (8 - c % 8);
        int64_t bytes = padded / 8;
        L->host_stride[lv] = (bytes + 7) / 8 * 8;
        L->host_base[lv] = base;
        base += 32 * L->host_stride[lv];
    }
    L->host_size = base;

    L->nchunks = 0;
    for (int lv = 0; lv < 4; ++lv) {
        L->chunks_per_plane[lv] = (L->host_stride[lv] + 16383) / 16384;
        L->nchunks += 32 * L->chunks_per_plane[lv];
    }
    L->trailer_fixed = 8 + 8 + 4 * 32 * 8;
    return PA_OK;
}

extern "C" int64_t pa_compress_bound(const pa_layout* L, int dtype) {
    if (!L || (dtype != PA_F32 && dtype != PA_F64)) return PA_ERR_ARG;
    return 8 + 2 * L->nchunks + L->host_size + 8 + L->trailer_fixed + L->n4 * dtype;
}

extern "C" const char* pa_version(void) { return "prism_amd 0.1.0 (gfx950)"; }
