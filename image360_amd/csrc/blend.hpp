// Internal interface of the chunk-blend kernels.
#pragma once

#include "common.hpp"

namespace i360 {

// SeamBlend::BlendChunkContributors on device-resident RGB48 contributors
// (`warped` is a host array of device pointers).  Enqueues on `s`; errors
// go to `st`.
void blend_chunk_u16(i360hip_ctx* ctx, const uint16_t* const* warped, int c, int w, int h, int num_bands,
                     uint16_t* out, int32_t* owner_out, hipStream_t s, Status& st);

// Frees the context's blend workspace.
void blend_release(i360hip_ctx* ctx);

// Same number of bands as SeamBlend.cpp:457-459 keeps.
inline int ClampBands(int w, int h, int num_bands) {
    int mb = 1;
    while (w > 1 && h > 1 && mb < num_bands) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
        ++mb;
    }
    int r = num_bands < mb ? num_bands : mb;
    return r < 1 ? 1 : r;
}

}  // namespace i360
