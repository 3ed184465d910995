// Float chunk render: internal descriptors shared by the planner and the
// kernels (render_float.hip).
#pragma once

#include "common.hpp"

namespace i360 {

constexpr int kMaxLevels = I360HIP_MAX_BANDS;

typedef _Float16 h4 __attribute__((ext_vector_type(4)));

// Storage/math precision of one render form.
enum RenderMode : int {
    kModeFP64 = I360HIP_RENDER_FP64,
    kModePlain = I360HIP_RENDER_PLAIN,
    kModePrecision = I360HIP_RENDER_PRECISION,
    kModeBatched = I360HIP_RENDER_BATCHED,
};

// One (chunk, contributor) pair.
struct RItem {
    const void* src;  // frame pixels, interleaved RGB (FP16 or FP32)
    int fw, fh, fmt;  // fmt: I360HIP_PIX_F16 / I360HIP_PIX_F32
    float gain;
    double hfwd[9];   // frame -> canvas
    double hinv[9];   // canvas -> frame (filled on the device, FP64)
    float hinv32[9];  // binary32 inverse of the binary32 forward matrix (plain form)
    int chunk;        // chunk index
    int local;        // index within the chunk's contributor list
    int wtile0, wtilen;  // this item's warp tiles (global tile indices)
    void* warped;     // chunk-sized RGBA (coverage in A), storage type of the form
    uint8_t* scov;    // coverage (0/1) in the chamfer's sheared plane, written by the warp over its
                      // footprint; read by the chamfer and the seam scan
    float* dist;      // sheared chamfer distances (inside the coverage box)
    int* box;         // coverage box x0,y0,x1,y1 (chunk-local, inclusive)
    int lbox[kMaxLevels][4];  // conservative support box per level (footprint dilated)
    void* nd[kMaxLevels];     // level l >= 1: RGBA (numerator rgb, denominator)
    void* wt[kMaxLevels];     // level l >= 1: weight
};

struct RChunk {
    int x, y, w, h;
    int first, count;  // items of this chunk
    int levels;
    int lw[kMaxLevels], lh[kMaxLevels];
    int* boxes;          // count * 4 (the items' box pointers point here)
    int8_t* owner;       // w*h contributor index (-1: none); the caller's owner_out is int32
    float* out;          // FP32 RGB, w*h*3 (caller's); RGB48 when out16
    int out16;           // out holds uint16 RGB48 (the reference's output rounding of v * 65535)
    int32_t* owner_out;  // optional (caller's)
    void* recon[kMaxLevels];  // level l >= 1 reconstruction, RGBA in the math type
    int* lohi;
    int* prevl;
    int* chosen;
    uint8_t* back;
    uint8_t* costs;  // seam band costs: 64 values of up to 8 bytes per line and slot
    int* counters;  // per-pair contested-line counts (count*count ints)
    // learned surrogate (SURROGATE form and its training-data export)
    int sk;           // pyramid level the surrogate predicts (levels-3); 0 = none for this chunk
    int spix0;        // first level-sk pixel of this chunk in the surrogate buffers
    int stile0, stx;  // first 16x16 surrogate tile of this chunk, tiles per row
};

// Surrogate buffers of a render (NHWC, indexed by RChunk::spix0 + pixel).
struct SurrBufs {
    float* feat;  // 16 features per level-K pixel
    float* aux;   // 4 floats per level-K pixel (first blur of the composite)
    float* out;   // 4 network outputs per level-K pixel
    float* tmp0;  // 32 floats per level-K pixel (network ping-pong)
    float* tmp1;
    float* target;  // export only: the plain blend's level-K reconstruction, 3 per pixel
    int* tflag;   // per 16x16 tile: fall back to the plain blend
    int* cflag;   // per chunk: some tile falls back (the deep levels are computed)
};

// A tile of some stage: id = item or chunk index, (tx, ty) tile coordinates.
struct RTile {
    int id, tx, ty, pad;
};

}  // namespace i360
