// GPU-resident float chunk render (i360hip_render_chunks): warp, coverage,
// chamfer ownership, DP seams, multi-band compose, for one chunk after
// another (FP64 / PLAIN / PRECISION) or for every chunk at once (BATCHED).
//
// Every stage is a table-driven kernel: the host planner lists the work
// items ((chunk, contributor) pairs), the chunks and, per stage, the tiles
// to process.  The per-chunk forms launch each stage on one chunk's slice
// of the tables; the batched form launches each stage once over all tiles
// of all chunks with a persistent grid and replays the launch sequence from
// a HIP graph.
//
// Support boxes: each work item's pixels outside its footprint (the
// frame's projected bounding box, widened by 2 px) are uncovered, so every
// level of its pyramids is exactly zero outside the footprint dilated by
// the blur support; kernels read such samples as 0 without touching memory
// and skip the contributor there (a zero weight adds nothing to a band).

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "blend.hpp"
#include <type_traits>

#include "chamfer.hpp"
#include "render.hpp"
#include "seam.hpp"
#include "surrogate.hpp"

namespace i360 {

namespace {

constexpr int kSlotRender = 12;
constexpr int kSlotRenderTables = 13;

// ------------------------------------------------------------ storage ---
template <int MODE> struct Form;
template <> struct Form<kModeFP64> { using M = double; using S = double; };
template <> struct Form<kModePlain> { using M = float; using S = float; };
template <> struct Form<kModePrecision> { using M = float; using S = _Float16; };

__device__ __forceinline__ void ld4(const void* p, size_t i, float (&o)[4], float*) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v q = gptr(static_cast<const f4v*>(p))[i];
    o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
}
__device__ __forceinline__ void ld4(const void* p, size_t i, float (&o)[4], _Float16*) {
    const h4 q = gptr(static_cast<const h4*>(p))[i];
    o[0] = static_cast<float>(q.x); o[1] = static_cast<float>(q.y);
    o[2] = static_cast<float>(q.z); o[3] = static_cast<float>(q.w);
}
__device__ __forceinline__ void ld4(const void* p, size_t i, double (&o)[4], double*) {
    typedef double d4v __attribute__((ext_vector_type(4)));
    const d4v q = gptr(static_cast<const d4v*>(p))[i];
    o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
}
__device__ __forceinline__ void st4(void* p, size_t i, const float (&v)[4], float*) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    f4v q;
    q.x = v[0]; q.y = v[1]; q.z = v[2]; q.w = v[3];
    gptr(static_cast<f4v*>(p))[i] = q;
}
__device__ __forceinline__ void st4(void* p, size_t i, const float (&v)[4], _Float16*) {
    h4 q;
    q.x = static_cast<_Float16>(v[0]); q.y = static_cast<_Float16>(v[1]);
    q.z = static_cast<_Float16>(v[2]); q.w = static_cast<_Float16>(v[3]);
    gptr(static_cast<h4*>(p))[i] = q;
}
__device__ __forceinline__ void st4(void* p, size_t i, const double (&v)[4], double*) {
    typedef double d4v __attribute__((ext_vector_type(4)));
    d4v q;
    q.x = v[0]; q.y = v[1]; q.z = v[2]; q.w = v[3];
    gptr(static_cast<d4v*>(p))[i] = q;
}
template <typename M, typename S>
__device__ __forceinline__ M ld1(const void* p, size_t i) {
    return static_cast<M>(gptr(static_cast<const S*>(p))[i]);
}
template <typename S, typename M>
__device__ __forceinline__ void st1(void* p, size_t i, M v) {
    gptr(static_cast<S*>(p))[i] = static_cast<S>(v);
}

// A box x0, y0, x1, y1 (inclusive) held in registers: read once per tile
// or item, so that no per-pixel test issues (and waits for) memory loads.
struct IBox {
    int x0, y0, x1, y1;
};
__device__ __forceinline__ IBox box_of(const int* b) {
    const auto* g = gptr(b);  // item boxes are 8-byte aligned: four dword loads
    return IBox{g[0], g[1], g[2], g[3]};
}
__device__ __forceinline__ bool in_box(const IBox& b, int x, int y) {
    return x >= b.x0 && x <= b.x1 && y >= b.y0 && y <= b.y1;
}

// ----------------------------------------------------------- FP64 map ---
__global__ void rmap_kernel(RItem* items, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* m = items[i].hfwd;
    const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], k = m[7], j = m[8];
    const double c0 = e * j - f * k, c1 = d * j - f * g, c2 = d * k - e * g;
    const double det = (a * c0 - b * c1) + c * c2;
    const double s = 1.0 / det;
    double* o = items[i].hinv;
    o[0] = c0 * s; o[1] = (c * k - b * j) * s; o[2] = (b * f - c * e) * s;
    o[3] = (f * g - d * j) * s; o[4] = (a * j - c * g) * s; o[5] = (c * d - a * f) * s;
    o[6] = c2 * s; o[7] = (b * g - a * k) * s; o[8] = (a * e - b * d) * s;
}

// Empty coverage boxes, and cleared contested-line counters of the pairs
// this item opens (row `local` of its chunk's counter matrix).
__global__ void rbox_init_kernel(RItem* items, const RChunk* chunks, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int* b = items[i].box;
    b[0] = INT_MAX; b[1] = INT_MAX; b[2] = -1; b[3] = -1;
    const RChunk& ch = chunks[items[i].chunk];
    for (int j = 0; j < ch.count; ++j) ch.counters[items[i].local * ch.count + j] = 0;
}

// --------------------------------------------------------------- warp ---
constexpr int kWarpTW = 64, kWarpTH = 16;

// RGB48 frames are read as v * (1/65535) in binary32, whatever the form's
// maths type (the value a binary32 [0,1] frame of the same image holds).
constexpr float kInv65535 = 1.0f / 65535.0f;

template <typename M>
__device__ __forceinline__ void fetch_rgb(const void* src, int fmt, size_t idx, M (&o)[3]) {
    if (fmt == I360HIP_PIX_F16) {
        const auto* p = gptr(static_cast<const _Float16*>(src) + idx * 3);
        o[0] = static_cast<M>(static_cast<float>(p[0]));
        o[1] = static_cast<M>(static_cast<float>(p[1]));
        o[2] = static_cast<M>(static_cast<float>(p[2]));
    } else if (fmt == I360HIP_PIX_U16) {
        const auto* p = gptr(static_cast<const uint16_t*>(src) + idx * 3);
        o[0] = static_cast<M>(static_cast<float>(p[0]) * kInv65535);
        o[1] = static_cast<M>(static_cast<float>(p[1]) * kInv65535);
        o[2] = static_cast<M>(static_cast<float>(p[2]) * kInv65535);
    } else {
        const auto* p = gptr(static_cast<const float*>(src) + idx * 3);
        o[0] = static_cast<M>(p[0]); o[1] = static_cast<M>(p[1]); o[2] = static_cast<M>(p[2]);
    }
}

typedef uint32_t u4a __attribute__((ext_vector_type(4), aligned(4)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

// Per-tile FP32 local map reduced from the FP64 inverse: for the tile-local
// offset d = (dx, dy), source = s0 + (A d) / (1 + g . d) exactly (the
// projective map re-centred on the tile origin), s0 split into an integer
// and a fraction so the binary32 part stays small.
struct TileMap {
    int ix, iy;
    float fx, fy, a00, a01, a10, a11, g0, g1, q0;
    float margin;  // bound on |FP32 position - FP64 position| (source px) inside the tile
    int ok;
    // every pixel of the tile lies in the chunk and the item's support box,
    // and its sample is valid with the FP32 position deciding it (exact
    // position at least twice the margin inside the frame, denominator
    // above 1/2): the warp may skip every per-pixel test
    int interior;
    // interior tiles: the frame texels [sx0, sx0+sw) x [sy0, sy0+sh) hold
    // every sample of the tile; sw > 0: the warp stages them in LDS
    int sx0, sy0, sw, sh;
};

// LDS staging of an interior warp tile's source texels: at most kStageRows
// frame rows of kStageWords 4-byte words (FP16 RGB, 6 bytes a texel,
// copied as 16-byte chunks from the 4-byte word at or below the box's
// first byte).
constexpr int kStageRows = 32;
constexpr int kStageWords = 128;

__device__ __forceinline__ TileMap tile_map(const double* h, double X0, double Y0) {
    TileMap t{};
    const double u0 = (h[0] * X0 + h[1] * Y0) + h[2];
    const double v0 = (h[3] * X0 + h[4] * Y0) + h[5];
    const double q0 = (h[6] * X0 + h[7] * Y0) + h[8];
    if (fabs(q0) < 1e-6) return t;  // ok = 0: per-pixel FP64 map for this tile
    const double sx = u0 / q0, sy = v0 / q0;
    if (fabs(sx) > 1e9 || fabs(sy) > 1e9) return t;
    const double fx = floor(sx), fy = floor(sy);
    t.ix = static_cast<int>(fx);
    t.iy = static_cast<int>(fy);
    t.fx = static_cast<float>(sx - fx);
    t.fy = static_cast<float>(sy - fy);
    t.a00 = static_cast<float>((h[0] - sx * h[6]) / q0);
    t.a01 = static_cast<float>((h[1] - sx * h[7]) / q0);
    t.a10 = static_cast<float>((h[3] - sy * h[6]) / q0);
    t.a11 = static_cast<float>((h[4] - sy * h[7]) / q0);
    t.g0 = static_cast<float>(h[6] / q0);
    t.g1 = static_cast<float>(h[7] / q0);
    t.q0 = static_cast<float>(q0);
    // position error: binary32 Jacobian terms (relative 2^-24 each) times
    // offsets <= 64, the division, and two roundings of the absolute
    // position; the 2e-3 px floor is ~100x the measured error
    const double jac = fabs(t.a00) + fabs(t.a01) + fabs(t.a10) + fabs(t.a11);
    const double mag = fmax(fabs(sx), fabs(sy)) + 64.0 * jac;
    t.margin = static_cast<float>(2e-3 + 64.0 * jac * 1.2e-6 + mag * 2.4e-7);
    t.ok = 1;
    return t;
}

// Tile-local FP32 source position of offset (dx, dy) (the precision
// forms' map where den > 1/2): both warp paths evaluate exactly this.
__device__ __forceinline__ void tile_pos(const TileMap& tm, int dx, int dy, float den, float& lx, float& ly) {
    const float r = __builtin_amdgcn_rcpf(den);  // 1 ulp; inside the tile's error margin
    lx = tm.fx + (tm.a00 * static_cast<float>(dx) + tm.a01 * static_cast<float>(dy)) * r;
    ly = tm.fy + (tm.a10 * static_cast<float>(dx) + tm.a11 * static_cast<float>(dy)) * r;
}
__device__ __forceinline__ float tile_den(const TileMap& tm, int dx, int dy) {
    return (1.0f + tm.g0 * static_cast<float>(dx)) + tm.g1 * static_cast<float>(dy);
}

// Validity of canvas pixel (x, y) of the chunk under the FP64 map
// (WarpPerspective_Scalar.cpp:108-122 in binary64).
__device__ __forceinline__ bool fp64_valid(const RItem& it, const RChunk& ch, int x, int y) {
    const double* m = it.hinv;
    const double X = static_cast<double>(ch.x + x), Y = static_cast<double>(ch.y + y);
    const double q = (m[6] * X + m[7] * Y) + m[8];
    if (fabs(q) < 1e-10) return false;
    const double u = ((m[0] * X + m[1] * Y) + m[2]) / q;
    const double v = ((m[3] * X + m[4] * Y) + m[5]) / q;
    const double fu = floor(u), fv = floor(v);
    return fu >= 0.0 && fv >= 0.0 && fu <= static_cast<double>(it.fw - 2) && fv <= static_cast<double>(it.fh - 2);
}

// Integer sample origin and fractions for one pixel; false if invalid.
template <int MODE, typename M>
__device__ __forceinline__ bool map_pixel(const RItem& it, const RChunk& ch, const TileMap& tm, int x, int y,
                                          int dx, int dy, int& sx0, int& sy0, M& fx, M& fy) {
    if (MODE == kModePrecision && tm.ok) {
        // The sample/no-sample decision is a discontinuity: take it from the
        // FP64 map, exactly as the FP64 form does, so both forms agree on
        // coverage (and hence on ownership).  The sample position comes from
        // the FP32 tile map; bilinear sampling is continuous across integer
        // boundaries, so a position that lands a hair past the last valid
        // cell is resolved to that cell's edge.
        //
        // Far from the horizon (den > 1/2, so |q| > 5e-7 and the FP64 map
        // is regular) the FP32 position decides whenever it is farther than
        // the tile's error bound from the validity edges (u = 0, u = fw-1,
        // v = 0, v = fh-1); only pixels within that band, a thin ring at the
        // footprint edge, evaluate the FP64 map.
        const float den = tile_den(tm, dx, dy);
        float lx, ly;
        if (den > 0.5f) {
            tile_pos(tm, dx, dy, den, lx, ly);
            const float ux = static_cast<float>(tm.ix) + lx, uy = static_cast<float>(tm.iy) + ly;
            const float ex = static_cast<float>(it.fw - 1), ey = static_cast<float>(it.fh - 1);
            const float m = tm.margin;
            if (fminf(fabsf(ux), fabsf(ux - ex)) < m || fminf(fabsf(uy), fabsf(uy - ey)) < m) {
                if (!fp64_valid(it, ch, x, y)) return false;
            } else if (!(ux >= 0.0f && ux < ex && uy >= 0.0f && uy < ey)) {
                return false;
            }
        } else {
            if (!fp64_valid(it, ch, x, y)) return false;
            lx = tm.fx + (tm.a00 * static_cast<float>(dx) + tm.a01 * static_cast<float>(dy)) / den;
            ly = tm.fy + (tm.a10 * static_cast<float>(dx) + tm.a11 * static_cast<float>(dy)) / den;
        }
        const float flx = floorf(lx), fly = floorf(ly);
        sx0 = tm.ix + static_cast<int>(flx);
        sy0 = tm.iy + static_cast<int>(fly);
        float ffx = lx - flx, ffy = ly - fly;
        if (sx0 > it.fw - 2) { sx0 = it.fw - 2; ffx = 1.0f; }
        if (sx0 < 0) { sx0 = 0; ffx = 0.0f; }
        if (sy0 > it.fh - 2) { sy0 = it.fh - 2; ffy = 1.0f; }
        if (sy0 < 0) { sy0 = 0; ffy = 0.0f; }
        fx = static_cast<M>(ffx);
        fy = static_cast<M>(ffy);
        return true;
    }
    if (MODE == kModePlain) {
        const float* m = it.hinv32;
        const float X = static_cast<float>(ch.x + x), Y = static_cast<float>(ch.y + y);
        float u = (m[0] * X + m[1] * Y) + m[2];
        float v = (m[3] * X + m[4] * Y) + m[5];
        const float q = (m[6] * X + m[7] * Y) + m[8];
        if (fabsf(q) < 1e-10f) return false;
        u = u / q;
        v = v / q;
        const float fu = floorf(u), fv = floorf(v);
        if (!(fu >= 0.0f && fv >= 0.0f && fu <= static_cast<float>(it.fw - 2) && fv <= static_cast<float>(it.fh - 2)))
            return false;
        sx0 = static_cast<int>(fu);
        sy0 = static_cast<int>(fv);
        fx = static_cast<M>(u - fu);
        fy = static_cast<M>(v - fv);
        return true;
    }
    // FP64 (also the precision form's fallback for degenerate tiles)
    const double* m = it.hinv;
    const double X = static_cast<double>(ch.x + x), Y = static_cast<double>(ch.y + y);
    double u = (m[0] * X + m[1] * Y) + m[2];
    double v = (m[3] * X + m[4] * Y) + m[5];
    const double q = (m[6] * X + m[7] * Y) + m[8];
    if (fabs(q) < 1e-10) return false;
    u = u / q;
    v = v / q;
    const double fu = floor(u), fv = floor(v);
    if (!(fu >= 0.0 && fv >= 0.0 && fu <= static_cast<double>(it.fw - 2) && fv <= static_cast<double>(it.fh - 2)))
        return false;
    sx0 = static_cast<int>(fu);
    sy0 = static_cast<int>(fv);
    fx = static_cast<M>(u - fu);
    fy = static_cast<M>(v - fv);
    return true;
}

// Raw 16 bytes holding texels idx, idx+1 of an FP16 RGB frame: the
// 4-byte-aligned window at or below byte 6*idx, clamped into the frame;
// `sh` receives the pair's byte offset in the window (0, 2, 4 or, for the
// frame's very last pair when its size is 2 mod 4, 6).
__device__ __forceinline__ u4a load_pair_f16(const void* src, size_t idx, size_t nbytes, uint32_t& sh) {
    const size_t o = idx * 6;
    const size_t o4 = min(o & ~static_cast<size_t>(3), (nbytes - 16) & ~static_cast<size_t>(3));
    sh = static_cast<uint32_t>(o - o4);
    // the window is only 4-byte aligned: address it directly (gptr's
    // parameter would carry the vector type's natural 16-byte alignment)
    return *(const I360_GLOBAL u4a*)(static_cast<const uint8_t*>(src) + o4);
}

template <typename M>
__device__ __forceinline__ void decode_pair_f16(const u4a& q, uint32_t sh, const void* src, size_t idx, M (&a)[3],
                                                M (&b)[3]) {
    if (sh > 4) {  // last pair of an odd-sized frame: texel by texel
        fetch_rgb<M>(src, I360HIP_PIX_F16, idx, a);
        fetch_rgb<M>(src, I360HIP_PIX_F16, idx + 1, b);
        return;
    }
    const bool s = sh >= 4;
    const uint32_t bits = (sh & 2u) * 8u;
    const uint32_t w0 = s ? q.y : q.x, w1 = s ? q.z : q.y, w2 = s ? q.w : q.z, w3 = s ? 0u : q.w;
    const h2v d0 = __builtin_bit_cast(h2v, __builtin_amdgcn_alignbit(w1, w0, bits));
    const h2v d1 = __builtin_bit_cast(h2v, __builtin_amdgcn_alignbit(w2, w1, bits));
    const h2v d2 = __builtin_bit_cast(h2v, __builtin_amdgcn_alignbit(w3, w2, bits));
    a[0] = static_cast<M>(static_cast<float>(d0.x));
    a[1] = static_cast<M>(static_cast<float>(d0.y));
    a[2] = static_cast<M>(static_cast<float>(d1.x));
    b[0] = static_cast<M>(static_cast<float>(d1.y));
    b[1] = static_cast<M>(static_cast<float>(d2.x));
    b[2] = static_cast<M>(static_cast<float>(d2.y));
}

// As decode_pair_f16, keeping the FP16 values (the precision forms blend
// them with mixed-precision FMAs, FP32 accumulation).
__device__ __forceinline__ void decode_pair_h(const u4a& q, uint32_t sh, const void* src, size_t idx,
                                              _Float16 (&a)[3], _Float16 (&b)[3]) {
    if (sh > 4) {  // last pair of an odd-sized frame: texel by texel
        const auto* p = gptr(static_cast<const _Float16*>(src) + idx * 3);
        a[0] = p[0]; a[1] = p[1]; a[2] = p[2];
        b[0] = p[3]; b[1] = p[4]; b[2] = p[5];
        return;
    }
    const bool s = sh >= 4;
    const uint32_t bits = (sh & 2u) * 8u;
    const uint32_t w0 = s ? q.y : q.x, w1 = s ? q.z : q.y, w2 = s ? q.w : q.z, w3 = s ? 0u : q.w;
    const h2v d0 = __builtin_bit_cast(h2v, __builtin_amdgcn_alignbit(w1, w0, bits));
    const h2v d1 = __builtin_bit_cast(h2v, __builtin_amdgcn_alignbit(w2, w1, bits));
    const h2v d2 = __builtin_bit_cast(h2v, __builtin_amdgcn_alignbit(w3, w2, bits));
    a[0] = d0.x; a[1] = d0.y; a[2] = d1.x;
    b[0] = d1.y; b[1] = d2.x; b[2] = d2.y;
}

// Whether every pixel of an FP16 item's warp tile lies in the chunk and the
// item's support box, has an FP32 denominator above 1/2, and has its exact
// source position at least twice the tile's error margin inside
// [0, fw-1) x [0, fh-1): then the FP32 position decides every pixel valid
// and no clamp changes it.  The FP32 denominator is monotone in dx and dy
// (so its minimum is at a corner), and with a positive denominator the
// projective image of the tile is the convex hull of its corners' images,
// so the four corners decide.
__device__ __forceinline__ bool tile_interior(const RItem& it, const RChunk& ch, const RTile& tl,
                                              const TileMap& tm) {
    if (!tm.ok || it.fmt != I360HIP_PIX_F16) return false;
    const int x0 = tl.tx * kWarpTW, y0 = tl.ty * kWarpTH;
    const int x1 = x0 + kWarpTW - 1, y1 = y0 + kWarpTH - 1;
    if (x1 >= ch.w || y1 >= ch.h) return false;
    const int* b = it.lbox[0];
    if (x0 < b[0] || y0 < b[1] || x1 > b[2] || y1 > b[3]) return false;
    const double m2 = 2.0 * static_cast<double>(tm.margin);
    const double ex = static_cast<double>(it.fw - 1) - m2, ey = static_cast<double>(it.fh - 1) - m2;
    const double* m = it.hinv;
    const double q0 = static_cast<double>(tm.q0);
    for (int k = 0; k < 4; ++k) {
        const int dx = (k & 1) ? kWarpTW - 1 : 0, dy = (k & 2) ? kWarpTH - 1 : 0;
        if (!(tile_den(tm, dx, dy) > 0.6f)) return false;
        const double X = static_cast<double>(ch.x + x0 + dx), Y = static_cast<double>(ch.y + y0 + dy);
        const double q = (m[6] * X + m[7] * Y) + m[8];
        if (!(q / q0 > 0.5)) return false;
        const double u = ((m[0] * X + m[1] * Y) + m[2]) / q;
        const double v = ((m[3] * X + m[4] * Y) + m[5]) / q;
        if (!(u >= m2 && u <= ex && v >= m2 && v <= ey)) return false;
    }
    return true;
}

// Tile-local first pixel (dx, dy) of the t-th (0..255) run of four pixels
// x, x+1, x+2, x+3 (x a multiple of 4) of a kWarpTW x kWarpTH tile, in the
// order that suits the chamfer's sheared planes.  Pixel (x, y) sits at
// sheared row y + (x >> shift), so the tile's 4-, 8- or 16-column groups
// form a parallelogram there; its (row, group) pairs are numbered by
// sheared row, then group, so a wavefront's 64 runs (64 dwords of a byte
// plane, 64 x 16 bytes of a float plane) lie in a few contiguous stretches
// of sheared rows instead of one stretch per column group.  The 256 runs
// are the tile's pixels, each once.
__device__ __forceinline__ void shear_tile_px(int t, int shift, int& dx, int& dy) {
    static_assert(kWarpTW == 64 && kWarpTH == 16, "tile shape of the numbering");
    const int S = 1 << shift, G = kWarpTW >> shift;  // column groups of the tile
    const int q = t >> (shift - 2), k = t & ((S >> 2) - 1);  // (row, group) pair, run in its group
    // largest d with d (d + 1) / 2 <= v
    auto tri = [](int v) {
        int d = static_cast<int>((sqrtf(8.0f * static_cast<float>(v) + 1.0f) - 1.0f) * 0.5f);
        if ((d + 1) * (d + 2) / 2 <= v) ++d;
        if (d * (d + 1) / 2 > v) --d;
        return d;
    };
    const int T1 = G * (G + 1) / 2, T3 = G * (G - 1) / 2;  // pairs on sheared rows d < G, d >= kWarpTH
    int d, gl;
    if (q < T1) {  // rows d < G: groups 0..d
        d = tri(q);
        gl = q - d * (d + 1) / 2;
    } else if (q < kWarpTH * G - T3) {  // rows G <= d < kWarpTH: all G groups
        const int q2 = q - T1;
        d = G + q2 / G;
        gl = q2 - (d - G) * G;
    } else {  // rows d >= kWarpTH: groups d - kWarpTH + 1 .. G - 1, counted from the end
        const int q3 = kWarpTH * G - 1 - q;
        const int dd = tri(q3);
        d = kWarpTH + G - 2 - dd;
        gl = G - 1 - (q3 - dd * (dd + 1) / 2);
    }
    dx = gl * S + 4 * k;
    dy = d - gl;
}

// Per-tile FP32 maps of the precision forms, one thread per warp tile
// (the FP64 reduction runs fully parallel here instead of serially inside
// the warp's tile loop).
__global__ void __launch_bounds__(256) rtilemap_kernel(const RItem* __restrict__ items,
                                                       const RChunk* __restrict__ chunks,
                                                       const RTile* __restrict__ tiles, int ntiles,
                                                       TileMap* __restrict__ maps) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ntiles) return;
    const RTile tl = tiles[e];
    const RItem& it = items[tl.id];
    const RChunk& ch = chunks[it.chunk];
    TileMap t = tile_map(it.hinv, static_cast<double>(ch.x + tl.tx * kWarpTW), static_cast<double>(ch.y + tl.ty * kWarpTH));
    t.interior = tile_interior(it, ch, tl, t) ? 1 : 0;
    if (t.interior) {
        // The exact positions of the tile's pixels lie in the convex hull of
        // its corners' (positive denominator), and every FP32 position is
        // within the margin of its exact one: the corners' FP32 box widened
        // by 1 + 2 * margin texels holds every pixel's (floor, floor + 1)
        // sample pair.  Staged only when every chunk load (16 bytes from the
        // word at or below a row's first byte, through the 4 words past its
        // last) stays inside the frame and the box fits the LDS tile.
        float mnx = 3e38f, mny = 3e38f, mxx = -3e38f, mxy = -3e38f;
        for (int k = 0; k < 4; ++k) {
            const int dx = (k & 1) ? kWarpTW - 1 : 0, dy = (k & 2) ? kWarpTH - 1 : 0;
            float px, py;
            tile_pos(t, dx, dy, tile_den(t, dx, dy), px, py);
            mnx = fminf(mnx, px); mny = fminf(mny, py);
            mxx = fmaxf(mxx, px); mxy = fmaxf(mxy, py);
        }
        const int pad = 1 + static_cast<int>(ceilf(2.0f * t.margin));
        const int bx0 = max(t.ix + static_cast<int>(floorf(mnx)) - pad, 0);
        const int by0 = max(t.iy + static_cast<int>(floorf(mny)) - pad, 0);
        const int bx1 = min(t.ix + static_cast<int>(floorf(mxx)) + 1 + pad, it.fw - 1);
        const int by1 = min(t.iy + static_cast<int>(floorf(mxy)) + 1 + pad, it.fh - 1);
        const long long w0 = (6ll * bx0) >> 2;
        const long long nw = ((6ll * (bx1 + 1) + 3) >> 2) - w0;
        const long long lastEnd = 6ll * by1 * it.fw + 4 * (w0 + 4 * ((nw + 3) >> 2));  // bytes
        if (bx1 >= bx0 && by1 >= by0 && by1 - by0 + 1 <= kStageRows && 4 * ((nw + 3) >> 2) <= kStageWords &&
            lastEnd <= 6ll * it.fw * it.fh) {
            t.sx0 = bx0; t.sy0 = by0; t.sw = bx1 - bx0 + 1; t.sh = by1 - by0 + 1;
        }
    }
    maps[e] = t;
}

// Warp of one item's tile.  Each wavefront owns 64x4 pixels and works
// alone (no LDS, no barrier): it maps its pixels (precision forms: from the
// tile's precomputed FP32 map), issues all 8 texel-row loads of a thread
// together (FP16 frames: unconditional, clamped addresses), blends,
// stores, and writes its partial coverage box (min/max over the wave by
// lane exchange) for rbox_reduce_kernel.
template <int MODE>
__global__ void __launch_bounds__(256) rwarp_kernel(const RItem* __restrict__ items,
                                                    const RChunk* __restrict__ chunks,
                                                    const RTile* __restrict__ tiles, int ntiles,
                                                    const TileMap* __restrict__ maps, int4* __restrict__ wbox,
                                                    int shift) {
    using M = typename Form<MODE == kModeBatched ? kModePrecision : MODE>::M;
    using S = typename Form<MODE == kModeBatched ? kModePrecision : MODE>::S;
    constexpr int MM = MODE == kModeBatched ? kModePrecision : MODE;
    constexpr int R = 4;
    const int wave = threadIdx.x >> 6;
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RItem& it = items[tl.id];
        const RChunk& ch = chunks[it.chunk];
        const int tx0 = tl.tx * kWarpTW, ty0 = tl.ty * kWarpTH;
        TileMap tm{};
        if (MM == kModePrecision) tm = maps[e];
        if constexpr (MM == kModePrecision) {
            if (tm.interior) {
                // every pixel valid, in the chunk and in the support box:
                // the same FP32 positions, texel loads and blend as below,
                // without the per-pixel tests, clamps and box reduction
                const void* src = it.src;
                const int fw = it.fw;
                const size_t nbytes = static_cast<size_t>(fw) * it.fh * 6;
                const float gain = it.gain;
                const bool useGain = fabsf(gain - 1.0f) >= 1e-6f;
                const int lx = threadIdx.x & 63;
                const int x = tx0 + lx;
                size_t i0[R];
                int sxa[R], sya[R];
                float fx[R], fy[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int ly = wave * R + r;
                    float px, py;
                    tile_pos(tm, lx, ly, tile_den(tm, lx, ly), px, py);
                    const float flx = floorf(px), fly = floorf(py);
                    sxa[r] = tm.ix + static_cast<int>(flx);
                    sya[r] = tm.iy + static_cast<int>(fly);
                    fx[r] = px - flx;
                    fy[r] = py - fly;
                    i0[r] = static_cast<size_t>(sya[r]) * fw + sxa[r];
                }
                // blend of row r's texel pairs (a, b) above (c, d), stores
                auto finish = [&](int r, float fxr, float fyr, const _Float16 (&a)[3], const _Float16 (&b)[3],
                                  const _Float16 (&c)[3], const _Float16 (&d)[3]) {
                    const float omx = 1.0f - fxr, omy = 1.0f - fyr;
                    const float wq[4] = {omx * omy, fxr * omy, omx * fyr, fxr * fyr};
                    float v[4];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        float acc = __builtin_fmaf(wq[0], static_cast<float>(a[k]), 0.0f);
                        acc = __builtin_fmaf(wq[1], static_cast<float>(b[k]), acc);
                        acc = __builtin_fmaf(wq[2], static_cast<float>(c[k]), acc);
                        float vv = __builtin_fmaf(wq[3], static_cast<float>(d[k]), acc);
                        if (useGain) vv = vv * gain;
                        v[k] = vv < 0.0f ? 0.0f : (1.0f < vv ? 1.0f : vv);
                    }
                    v[3] = 1.0f;
                    const int y = ty0 + wave * R + r;
                    const size_t p = static_cast<size_t>(y) * ch.w + x;
                    st4(it.warped, p, v, static_cast<S*>(nullptr));
                };
                if (tm.sw > 0) {  // tile-uniform
                    // Stage the tile's source box: half a wavefront per box
                    // row, 16-byte chunks from the word at or below the
                    // row's first byte (every chunk inside the frame: the
                    // tile map checked the last one).
                    __shared__ __attribute__((aligned(16))) uint32_t stage[kStageRows * kStageWords + 4];
                    const auto* src8 = static_cast<const uint8_t*>(src);
                    const size_t rowBytes = static_cast<size_t>(fw) * 6;
                    const int w0 = (6 * tm.sx0) >> 2;
                    const int nc = (((6 * (tm.sx0 + tm.sw) + 3) >> 2) - w0 + 3) >> 2;  // chunks per row, <= 32
                    const int hl = lx >> 5, cl = lx & 31;
                    u4a ldv[kStageRows / 8];
#pragma unroll
                    for (int k = 0; k < kStageRows / 8; ++k) {
                        const int r = 8 * k + 2 * wave + hl;
                        const bool act = r < tm.sh && cl < nc;
                        const size_t byte = static_cast<size_t>(tm.sy0 + (act ? r : 0)) * rowBytes +
                                            4 * static_cast<size_t>(w0 + (act ? 4 * cl : 0));
                        ldv[k] = *(const I360_GLOBAL u4a*)(src8 + byte);
                    }
                    lds_barrier();  // the previous staged tile's samples are read
#pragma unroll
                    for (int k = 0; k < kStageRows / 8; ++k) {
                        const int r = 8 * k + 2 * wave + hl;
                        if (r < tm.sh && cl < nc) {
                            typedef uint32_t u4s __attribute__((ext_vector_type(4)));
                            u4s q;
                            q.x = ldv[k].x; q.y = ldv[k].y; q.z = ldv[k].z; q.w = ldv[k].w;
                            *reinterpret_cast<u4s*>(stage + r * kStageWords + 4 * cl) = q;
                        }
                    }
                    // the staged words count as read here on every path (a
                    // thread with no row to write reads none): their loads
                    // are then known complete when the registers are reused
#pragma unroll
                    for (int k = 0; k < kStageRows / 8; ++k)
                        asm volatile("" ::"v"(ldv[k].x), "v"(ldv[k].y), "v"(ldv[k].z), "v"(ldv[k].w));
                    lds_barrier();
                    // Rows whose pair lies in the staged box blend from LDS;
                    // the others (rounding beyond the margin, rare) are done
                    // afterwards from the frame, their position recomputed.
                    // Kept apart, the rare rows' global loads share no
                    // registers with the staged rows, whose LDS reads then
                    // wait for no memory operation (they had waited for
                    // every store before).
                    uint32_t farm = 0u;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int lr = sya[r] - tm.sy0, lc = sxa[r] - tm.sx0;
                        if (lr >= 0 && lr + 1 < tm.sh && lc >= 0 && lc + 1 < tm.sw) {
                            _Float16 a[3], b[3], c[3], d[3];
                            // the pair's 12 bytes start 0 or 2 bytes into word q
                            // (so the decode's odd-frame path, which reads the
                            // frame at the index, is never taken here)
                            const int off = 6 * sxa[r] - 4 * w0;
                            const uint32_t* t0 = stage + lr * kStageWords + (off >> 2);
                            const uint32_t* t1 = t0 + kStageWords;
                            const u4a q0 = {t0[0], t0[1], t0[2], t0[3]};
                            const u4a q1 = {t1[0], t1[1], t1[2], t1[3]};
                            decode_pair_h(q0, static_cast<uint32_t>(off & 3), src, 0, a, b);
                            decode_pair_h(q1, static_cast<uint32_t>(off & 3), src, 0, c, d);
                            finish(r, fx[r], fy[r], a, b, c, d);
                        } else {
                            farm |= 1u << r;
                        }
                    }
                    if (__any(farm != 0u)) {  // wave-uniform
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            if (!((farm >> r) & 1u)) continue;
                            const int ly = wave * R + r;
                            float px, py;
                            tile_pos(tm, lx, ly, tile_den(tm, lx, ly), px, py);
                            const float flx = floorf(px), fly = floorf(py);
                            const size_t ir = static_cast<size_t>(tm.iy + static_cast<int>(fly)) * fw +
                                              (tm.ix + static_cast<int>(flx));
                            _Float16 a[3], b[3], c[3], d[3];
                            uint32_t s0, s1;
                            const u4a q0 = load_pair_f16(src, ir, nbytes, s0);
                            const u4a q1 = load_pair_f16(src, ir + fw, nbytes, s1);
                            decode_pair_h(q0, s0, src, ir, a, b);
                            decode_pair_h(q1, s1, src, ir + fw, c, d);
                            finish(r, px - flx, py - fly, a, b, c, d);
                        }
                    }
                } else {
                    u4a q0[R], q1[R];
                    uint32_t s0[R], s1[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        q0[r] = load_pair_f16(src, i0[r], nbytes, s0[r]);
                        q1[r] = load_pair_f16(src, i0[r] + fw, nbytes, s1[r]);
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        _Float16 a[3], b[3], c[3], d[3];
                        decode_pair_h(q0[r], s0[r], src, i0[r], a, b);
                        decode_pair_h(q1[r], s1[r], src, i0[r] + fw, c, d);
                        finish(r, fx[r], fy[r], a, b, c, d);
                    }
                }
                // the tile's coverage (all 1): one dword per thread, runs
                // along the sheared rows
                {
                    int dx, dy;
                    shear_tile_px(static_cast<int>(threadIdx.x), shift, dx, dy);
                    *gptr(reinterpret_cast<uint32_t*>(it.scov + ShearAt(tx0 + dx, ty0 + dy, shift, DistPitch(ch.w)))) =
                        0x01010101u;
                }
                if (lx == 0) {
                    typedef int i4v __attribute__((ext_vector_type(4)));
                    i4v bv;
                    bv.x = tx0; bv.y = ty0 + wave * R; bv.z = tx0 + kWarpTW - 1; bv.w = ty0 + wave * R + R - 1;
                    *gptr(reinterpret_cast<i4v*>(wbox) + static_cast<size_t>(e) * 4 + wave) = bv;
                }
                continue;
            }
        }
        const IBox fb = box_of(it.lbox[0]);
        const void* src = it.src;
        const int fw = it.fw, fmt = it.fmt;
        const size_t nbytes = static_cast<size_t>(fw) * it.fh * (fmt == I360HIP_PIX_F16 ? 6 : 12);
        const float gain = it.gain;
        const bool useGain = fabsf(gain - 1.0f) >= 1e-6f;
        const int lx = threadIdx.x & 63;
        const int x = tx0 + lx;
        // phase 1: map
        bool in[R], ok[R];
        size_t i0[R];
        M fx[R], fy[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int ly = wave * R + r;
            const int y = ty0 + ly;
            in[r] = x < ch.w && y < ch.h && in_box(fb, x, y);
            int sx0 = 0, sy0 = 0;
            fx[r] = M(0);
            fy[r] = M(0);
            ok[r] = in[r] && map_pixel<MM, M>(it, ch, tm, x, y, lx, ly, sx0, sy0, fx[r], fy[r]);
            i0[r] = ok[r] ? static_cast<size_t>(sy0) * fw + sx0 : 0;
        }
        // phase 2: loads and the bilinear blend of every row
        M val[R][3];
        auto weights = [&](int r, M (&wq)[4]) {
            const M omx = M(1) - fx[r], omy = M(1) - fy[r];
            wq[0] = omx * omy; wq[1] = fx[r] * omy; wq[2] = omx * fy[r]; wq[3] = fx[r] * fy[r];
        };
        if (fmt == I360HIP_PIX_F16) {
            u4a q0[R], q1[R];
            uint32_t s0[R], s1[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                q0[r] = load_pair_f16(src, i0[r], nbytes, s0[r]);
                q1[r] = load_pair_f16(src, i0[r] + fw, nbytes, s1[r]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                M wq[4];
                weights(r, wq);
                if constexpr (MM == kModePrecision) {
                    // precision forms: the FP16 texels go straight into FP32
                    // FMAs (mixed-precision multiply-add), same order of terms
                    _Float16 a[3], b[3], c[3], d[3];
                    decode_pair_h(q0[r], s0[r], src, i0[r], a, b);
                    decode_pair_h(q1[r], s1[r], src, i0[r] + fw, c, d);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        float acc = __builtin_fmaf(wq[0], static_cast<float>(a[k]), 0.0f);
                        acc = __builtin_fmaf(wq[1], static_cast<float>(b[k]), acc);
                        acc = __builtin_fmaf(wq[2], static_cast<float>(c[k]), acc);
                        val[r][k] = __builtin_fmaf(wq[3], static_cast<float>(d[k]), acc);
                    }
                } else {
                    M a[3], b[3], c[3], d[3];
                    decode_pair_f16<M>(q0[r], s0[r], src, i0[r], a, b);
                    decode_pair_f16<M>(q1[r], s1[r], src, i0[r] + fw, c, d);
#pragma unroll
                    for (int k = 0; k < 3; ++k) val[r][k] = ((wq[0] * a[k] + wq[1] * b[k]) + wq[2] * c[k]) + wq[3] * d[k];
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                M a[3], b[3], c[3], d[3];
                fetch_rgb<M>(src, fmt, i0[r], a);
                fetch_rgb<M>(src, fmt, i0[r] + 1, b);
                fetch_rgb<M>(src, fmt, i0[r] + fw, c);
                fetch_rgb<M>(src, fmt, i0[r] + fw + 1, d);
                M wq[4];
                weights(r, wq);
#pragma unroll
                for (int k = 0; k < 3; ++k) val[r][k] = ((wq[0] * a[k] + wq[1] * b[k]) + wq[2] * c[k]) + wq[3] * d[k];
            }
        }
        // phase 3: gain, clamp, store
        int bx0 = INT_MAX, by0 = INT_MAX, bx1 = -1, by1 = -1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (!in[r]) continue;
            const int y = ty0 + wave * R + r;
            const size_t p = static_cast<size_t>(y) * ch.w + x;
            M v[4] = {M(0), M(0), M(0), M(0)};
            if (ok[r]) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    M vv = val[r][k];
                    if (useGain) vv = vv * static_cast<M>(gain);
                    v[k] = vv < M(0) ? M(0) : (M(1) < vv ? M(1) : vv);
                }
                v[3] = M(1);
                bx0 = min(bx0, x); by0 = min(by0, y); bx1 = max(bx1, x); by1 = max(by1, y);
            }
            st4(it.warped, p, v, static_cast<S*>(nullptr));
            // coverage, for the chamfer and the seam scan
            gptr(it.scov)[ShearAt(x, y, shift, DistPitch(ch.w))] = ok[r] ? 1 : 0;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            bx0 = min(bx0, __shfl_xor(bx0, o));
            by0 = min(by0, __shfl_xor(by0, o));
            bx1 = max(bx1, __shfl_xor(bx1, o));
            by1 = max(by1, __shfl_xor(by1, o));
        }
        if (lx == 0) {
            typedef int i4v __attribute__((ext_vector_type(4)));
            i4v v;
            v.x = bx0; v.y = by0; v.z = bx1; v.w = by1;
            *gptr(reinterpret_cast<i4v*>(wbox) + static_cast<size_t>(e) * 4 + wave) = v;
        }
    }
}

// Coverage box of each item from its warp tiles' per-wave partial boxes
// (one block per item).
__global__ void __launch_bounds__(256) rbox_reduce_kernel(RItem* __restrict__ items, int n,
                                                          const int4* __restrict__ wbox) {
    __shared__ int red[4][256];
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const RItem& it = items[i];
        int b0 = INT_MAX, b1 = INT_MAX, b2 = -1, b3 = -1;
        const size_t q0 = static_cast<size_t>(it.wtile0) * 4, q1 = q0 + static_cast<size_t>(it.wtilen) * 4;
        for (size_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
            const int4 v = wbox[q];
            b0 = min(b0, v.x); b1 = min(b1, v.y); b2 = max(b2, v.z); b3 = max(b3, v.w);
        }
        red[0][threadIdx.x] = b0; red[1][threadIdx.x] = b1; red[2][threadIdx.x] = b2; red[3][threadIdx.x] = b3;
        __syncthreads();
        for (int k = 128; k > 0; k >>= 1) {
            if (threadIdx.x < k) {
                red[0][threadIdx.x] = min(red[0][threadIdx.x], red[0][threadIdx.x + k]);
                red[1][threadIdx.x] = min(red[1][threadIdx.x], red[1][threadIdx.x + k]);
                red[2][threadIdx.x] = max(red[2][threadIdx.x], red[2][threadIdx.x + k]);
                red[3][threadIdx.x] = max(red[3][threadIdx.x], red[3][threadIdx.x + k]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            int* b = it.box;
            if (red[2][0] >= 0) { b[0] = red[0][0]; b[1] = red[1][0]; b[2] = red[2][0]; b[3] = red[3][0]; }
            else { b[0] = INT_MAX; b[1] = INT_MAX; b[2] = -1; b[3] = -1; }
        }
        __syncthreads();
    }
}

// Largest |FP32 tile map - FP64 map| in source pixels (diagnostic).
__global__ void __launch_bounds__(256) rmap_error_kernel(const RItem* __restrict__ items,
                                                         const RChunk* __restrict__ chunks,
                                                         const RTile* __restrict__ tiles, int ntiles,
                                                         unsigned int* __restrict__ maxbits) {
    __shared__ TileMap smap;
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RItem& it = items[tl.id];
        const RChunk& ch = chunks[it.chunk];
        const int tx0 = tl.tx * kWarpTW, ty0 = tl.ty * kWarpTH;
        if (threadIdx.x == 0)
            smap = tile_map(it.hinv, static_cast<double>(ch.x + tx0), static_cast<double>(ch.y + ty0));
        __syncthreads();
        const TileMap tm = smap;
        float worst = 0.0f;
        const int lx = threadIdx.x & 63;
        for (int r = 0; r < 4; ++r) {
            const int ly = (threadIdx.x >> 6) * 4 + r;
            const int x = tx0 + lx, y = ty0 + ly;
            if (x >= ch.w || y >= ch.h || !tm.ok) continue;
            const double* m = it.hinv;
            const double X = ch.x + x, Y = ch.y + y;
            const double q = (m[6] * X + m[7] * Y) + m[8];
            const double u = ((m[0] * X + m[1] * Y) + m[2]) / q, v = ((m[3] * X + m[4] * Y) + m[5]) / q;
            if (!(u >= 0.0 && v >= 0.0 && u < it.fw - 1 && v < it.fh - 1)) continue;
            const float den = (1.0f + tm.g0 * static_cast<float>(lx)) + tm.g1 * static_cast<float>(ly);
            float ox, oy;  // as map_pixel computes them
            if (den > 0.5f) {
                const float rd = __builtin_amdgcn_rcpf(den);  // as the warp evaluates it
                ox = (tm.a00 * static_cast<float>(lx) + tm.a01 * static_cast<float>(ly)) * rd;
                oy = (tm.a10 * static_cast<float>(lx) + tm.a11 * static_cast<float>(ly)) * rd;
            } else {
                ox = (tm.a00 * static_cast<float>(lx) + tm.a01 * static_cast<float>(ly)) / den;
                oy = (tm.a10 * static_cast<float>(lx) + tm.a11 * static_cast<float>(ly)) / den;
            }
            const double eu = (tm.ix + static_cast<double>(tm.fx + ox)) - u;
            const double ev = (tm.iy + static_cast<double>(tm.fy + oy)) - v;
            worst = fmaxf(worst, static_cast<float>(fmax(fabs(eu), fabs(ev))));
        }
        atomicMax(maxbits, __float_as_uint(worst));
        __syncthreads();
    }
}

// ---------------------------------------------------------- ownership ---
// Owner = the contributor with the largest distance, the first on ties
// (distances are read from the chamfer's sheared planes, shift = the
// launch's ChamferShear).  A 256-thread block takes one 64x16 tile per
// iteration; the chunk's distance planes and coverage boxes are staged in
// LDS once per tile (double buffered by iteration, one barrier per tile).
// Every contributor's distance load is issued unconditionally (a run
// outside the box or the chunk reads its plane's first elements, and the
// values are discarded), so the loads of several contributors are in
// flight together.
struct OwnerItem {
    const float* dist;
    int box[4];
};
constexpr int kOwnerPix = 4;

// Ownership of one 64x16 tile of a chunk: argmax over the contributors of
// the chamfer distance, > 0, ties to the lowest index (SeamBlend.cpp:244-261).
// Each thread owns a run of four pixels in the sheared-row order of
// shear_tile_px: its distances are one 16-byte load per contributor from
// the sheared plane, and a wavefront's loads lie in a few stretches of
// sheared rows.  The owner map is written as bytes (a dword when the row
// pitch keeps the run 4-byte aligned).
__global__ void __launch_bounds__(kWarpTW * kWarpTH / kOwnerPix) rowner_kernel(const RItem* __restrict__ items,
                                                                              const RChunk* __restrict__ chunks,
                                                                              const RTile* __restrict__ tiles,
                                                                              int ntiles, int shift) {
    static_assert(kOwnerPix == 4 && kWarpTW * kWarpTH / kOwnerPix == 256, "one four-pixel run per thread");
    __shared__ OwnerItem sit[2][I360HIP_MAX_CONTRIBUTORS];
    int dx, dy;
    shear_tile_px(static_cast<int>(threadIdx.x), shift, dx, dy);
    int buf = 0;
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x, buf ^= 1) {
        const RTile tl = tiles[e];
        const RChunk& ch = chunks[tl.id];
        const int w = ch.w, h = ch.h, count = ch.count;
        if (static_cast<int>(threadIdx.x) < count) {
            const RItem& it = items[ch.first + threadIdx.x];
            const auto* b = gptr(it.box);
            OwnerItem f;
            f.dist = it.dist;
            f.box[0] = b[0]; f.box[1] = b[1]; f.box[2] = b[2]; f.box[3] = b[3];
            sit[buf][threadIdx.x] = f;
        }
        lds_barrier();
        const int x = tl.tx * kWarpTW + dx, y = tl.ty * kWarpTH + dy;
        const bool live = x < w && y < h;  // the run's first pixel exists (x < w: the run lies in the pitch)
        const size_t so = live ? ShearAt(x, y, shift, DistPitch(w)) : 0;
        int best[4] = {-1, -1, -1, -1};
        float bd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
        for (int i = 0; i < count; ++i) {
            const OwnerItem it = sit[buf][i];  // one LDS read; the box test below is branch-free
            // bit k: pixel x + k exists and lies in the item's box
            // (lo in [0, 4], hi in [-1, 3]; the mask is empty when lo > hi)
            const int lo = min(max(x, it.box[0]) - x, 4), hi = max(min(min(x + 3, it.box[2]), w - 1) - x, -1);
            const bool rowIn = live & (y >= it.box[1]) & (y <= it.box[3]);
            const uint32_t in = rowIn ? (((1u << (hi + 1)) - 1u) & ~((1u << lo) - 1u)) : 0u;
            const f4v v = ld_f4(it.dist + (in ? so : 0));  // distances exist inside the box only
            const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (((in >> k) & 1u) && vv[k] > bd[k]) { bd[k] = vv[k]; best[k] = i; }
        }
        if (live) {
            int8_t* o = ch.owner + static_cast<size_t>(y) * w + x;
            if ((w & 3) == 0) {  // tile-uniform; the run is whole and 4-byte aligned
                const uint32_t q = (static_cast<uint32_t>(best[0]) & 0xffu) | ((static_cast<uint32_t>(best[1]) & 0xffu) << 8) |
                                   ((static_cast<uint32_t>(best[2]) & 0xffu) << 16) |
                                   ((static_cast<uint32_t>(best[3]) & 0xffu) << 24);
                *gptr(reinterpret_cast<uint32_t*>(o)) = q;
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (x + k < w) gptr(o)[k] = static_cast<int8_t>(best[k]);
            }
        }
    }
}

// float-valued contributor image for the seam DP: RGBA storage -> M
template <typename S, typename M>
struct RgbaImg {
    const void* p;
    __device__ __forceinline__ void rgb(size_t i, M (&o)[3]) const {
        M v[4];
        ld4(p, i, v, static_cast<S*>(nullptr));
        o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
    }
    __device__ static constexpr M smooth() { return M(2) / M(65535); }
};

// Contributor image for the seam DP of the precision forms: each sample is
// recomputed from the source frame with the FP64 map and binary64
// bilinear weights (the FP64 form's warp, value for value), so the DP
// costs, and therefore the seams, equal the FP64 form's.
struct MapImg {
    const void* src;
    int fw, fh, fmt;
    float gain;
    double h[9];
    int cx, cy, w;
    __device__ __forceinline__ void rgb(size_t i, double (&o)[3]) const {
        // chunk pixel indices are below 2^31 (the planner rejects larger
        // chunks): a 32-bit division
        const uint32_t ii = static_cast<uint32_t>(i), uw = static_cast<uint32_t>(w);
        const uint32_t yq = ii / uw;
        const int x = static_cast<int>(ii - yq * uw), y = static_cast<int>(yq);
        const double X = static_cast<double>(cx + x), Y = static_cast<double>(cy + y);
        double u = (h[0] * X + h[1] * Y) + h[2];
        double v = (h[3] * X + h[4] * Y) + h[5];
        const double q = (h[6] * X + h[7] * Y) + h[8];
        o[0] = o[1] = o[2] = 0.0;
        if (fabs(q) < 1e-10) return;
        u = u / q;
        v = v / q;
        const double fu = floor(u), fv = floor(v);
        if (!(fu >= 0.0 && fv >= 0.0 && fu <= static_cast<double>(fw - 2) && fv <= static_cast<double>(fh - 2))) return;
        const int sx0 = static_cast<int>(fu), sy0 = static_cast<int>(fv);
        const double fx = u - fu, fy = v - fv;
        const double omx = 1.0 - fx, omy = 1.0 - fy;
        const double w00 = omx * omy, w10 = fx * omy, w01 = omx * fy, w11 = fx * fy;
        double a[3], b[3], c[3], d[3];
        const size_t i0 = static_cast<size_t>(sy0) * fw + sx0;
        fetch_rgb<double>(src, fmt, i0, a);
        fetch_rgb<double>(src, fmt, i0 + 1, b);
        fetch_rgb<double>(src, fmt, i0 + fw, c);
        fetch_rgb<double>(src, fmt, i0 + fw + 1, d);
        const bool useGain = fabsf(gain - 1.0f) >= 1e-6f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double val = ((w00 * a[k] + w10 * b[k]) + w01 * c[k]) + w11 * d[k];
            if (useGain) val = val * static_cast<double>(gain);
            o[k] = val < 0.0 ? 0.0 : (1.0 < val ? 1.0 : val);
        }
    }
    __device__ static constexpr double smooth() { return 2.0 / 65535.0; }
};

// ------------------------------------------------------------ reduce ---
constexpr int kRedTW = 64, kRedTH = 8, kRedRows = 2 * kRedTH + 3;

template <typename M>
__device__ __forceinline__ M blur5m(M a, M b, M m, M d, M e) {
    const M k0 = M(1) / M(16), k1 = M(4) / M(16), k2 = M(6) / M(16);
    const M ta = k0 * a, tb = k1 * b, tm = k2 * m, td = k1 * d, te = k0 * e;
    return (((ta + tb) + tm) + td) + te;
}

// The same 5-tap binomial blur in symmetric form, k0 (a + e) + k1 (b + d) +
// k2 m with two fused multiply-adds (5 instructions instead of 9): the
// FP16-storage forms' reduce (FP32 arithmetic; its rounding differs from
// blur5m's within FP32, far inside the FP16 storage error).
__device__ __forceinline__ float blur5s(float a, float b, float m, float d, float e) {
    return __builtin_fmaf(1.0f / 16.0f, a + e, __builtin_fmaf(4.0f / 16.0f, b + d, (6.0f / 16.0f) * m));
}

// One tile of level l+1 for one item: level-l samples outside the item's
// support read as zero.
// Surrogate phases of the pyramid stages: 0 = every level; 1 = the levels
// up to the surrogate level K of each chunk; 2 = the levels above K, only
// for chunks with a fallback tile (cflag).  Chunks without a surrogate
// level do everything in phase 1.
__device__ __forceinline__ bool phase_skip(const RChunk& ch, int c, int lvl, int phase, const int* cflag) {
    if (phase == 0) return false;
    const int K = ch.sk >= 1 ? ch.sk : kMaxLevels;
    if (phase == 1) return lvl > K;
    return lvl <= K || !cflag[c];
}

template <int MODE, bool L0>
__global__ void __launch_bounds__(256) rreduce_kernel(const RItem* __restrict__ items,
                                                      const RChunk* __restrict__ chunks,
                                                      const RTile* __restrict__ tiles, int ntiles, int l,
                                                      int phase, const int* __restrict__ cflag) {
    constexpr int MM = MODE == kModeBatched ? kModePrecision : MODE;
    using M = typename Form<MM>::M;
    using S = typename Form<MM>::S;
    __shared__ M tmp[5][kRedRows][kRedTW];
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RItem& it = items[tl.id];
        const RChunk& ch = chunks[it.chunk];
        if (phase_skip(ch, it.chunk, l + 1, phase, cflag)) continue;  // block-uniform
        const int w = ch.lw[l], h = ch.lh[l], ow = ch.lw[l + 1], oh = ch.lh[l + 1];
        const int ox0 = tl.tx * kRedTW, oy0 = tl.ty * kRedTH;
        const IBox sb = box_of(L0 ? it.box : it.lbox[l]);
        for (int idx = threadIdx.x; idx < kRedRows * kRedTW; idx += 256) {
            const int rr = idx / kRedTW, xx = idx % kRedTW;
            const int ox = ox0 + xx;
            if (ox >= ow) continue;
            const int sx = min(2 * ox, w - 1);
            const int ry = min(max(2 * oy0 - 2 + rr, 0), h - 1);
            M v[5][5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const int cx = min(max(sx - 2 + k, 0), w - 1);
#pragma unroll
                for (int c = 0; c < 5; ++c) v[k][c] = M(0);
                if (in_box(sb, cx, ry)) {
                    const size_t p = static_cast<size_t>(ry) * w + cx;
                    M q[4];
                    if (L0) {
                        ld4(it.warped, p, q, static_cast<S*>(nullptr));
                        v[k][0] = q[0] * q[3];
                        v[k][1] = q[1] * q[3];
                        v[k][2] = q[2] * q[3];
                        v[k][3] = q[3];
                        v[k][4] = gptr(ch.owner)[p] == it.local ? M(1) : M(0);
                    } else {
                        ld4(it.nd[l], p, q, static_cast<S*>(nullptr));
                        v[k][0] = q[0]; v[k][1] = q[1]; v[k][2] = q[2]; v[k][3] = q[3];
                        v[k][4] = ld1<M, S>(it.wt[l], p);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) tmp[c][rr][xx] = blur5m<M>(v[0][c], v[1][c], v[2][c], v[3][c], v[4][c]);
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < kRedTH * kRedTW; idx += 256) {
            const int yy = idx / kRedTW, xx = idx % kRedTW;
            const int ox = ox0 + xx, oy = oy0 + yy;
            if (ox >= ow || oy >= oh) continue;
            const int r = 2 * yy + 2;
            M o[5];
#pragma unroll
            for (int c = 0; c < 5; ++c)
                o[c] = blur5m<M>(tmp[c][r - 2][xx], tmp[c][r - 1][xx], tmp[c][r][xx], tmp[c][r + 1][xx], tmp[c][r + 2][xx]);
            const size_t q = static_cast<size_t>(oy) * ow + ox;
            const M nd[4] = {o[0], o[1], o[2], o[3]};
            st4(it.nd[l + 1], q, nd, static_cast<S*>(nullptr));
            st1<S, M>(it.wt[l + 1], q, o[4]);
        }
        __syncthreads();
    }
}

// Level l -> l+1 of the FP16 forms: the tile's level-l window (19 rows x
// 132 columns, replicate-clamped, zero outside the item's support) is
// staged into LDS, and both blur passes read LDS.  Same arithmetic as
// rreduce_kernel<MODE, L0>.  Level 0 reads the warped plane (premultiplied
// by its coverage) and the owner map (weight = owned by this item); the
// levels above read the numerator/denominator and weight planes.
// Software pipelined: the next tile's window loads are issued as soon as
// the current window is in LDS, and stay in flight across the blur passes
// (the barriers order LDS only).
constexpr int kRed0Cols = 2 * kRedTW + 4;
template <int MODE, bool L0>
__global__ void __launch_bounds__(256) rreduce_s_kernel(const RItem* __restrict__ items,
                                                        const RChunk* __restrict__ chunks,
                                                        const RTile* __restrict__ tiles, int ntiles, int l,
                                                        int phase, const int* __restrict__ cflag) {
    using M = float;
    constexpr int NE = kRedRows * kRed0Cols;  // staged elements
    constexpr int PER = (NE + 255) / 256;
    __shared__ h4 sq[NE];
    __shared__ _Float16 sw[NE];
    __shared__ M tmp[5][kRedRows][kRedTW];
    // first tile at or after f (stride gridDim.x) that this phase processes
    auto next_tile = [&](int f) {
        for (; f < ntiles; f += gridDim.x) {
            const RItem& it = items[tiles[f].id];
            if (!phase_skip(chunks[it.chunk], it.chunk, l + 1, phase, cflag)) break;
        }
        return f;
    };
    h4 qv[PER];
    int ov[PER];       // level 0: owner
    _Float16 wv[PER];  // levels above: weight
    bool inb[PER];
    // issues tile f's window loads (every address valid: elements outside
    // the support read element 0 and are zeroed at the commit)
    auto fetch = [&](int f) {
        const RTile tl = tiles[f];
        const RItem& it = items[tl.id];
        const RChunk& ch = chunks[it.chunk];
        const int w = ch.lw[l], h = ch.lh[l];
        const int ox0 = tl.tx * kRedTW, oy0 = tl.ty * kRedTH;
        const IBox sb = box_of(L0 ? it.box : it.lbox[l]);
        const h4* qp = static_cast<const h4*>(L0 ? it.warped : it.nd[l]);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + 256 * k;
            const int rr = i / kRed0Cols, j = i % kRed0Cols;
            const int cx = min(max(2 * ox0 - 2 + j, 0), w - 1);
            const int ry = min(max(2 * oy0 - 2 + rr, 0), h - 1);
            inb[k] = i < NE && in_box(sb, cx, ry);
            const size_t p = inb[k] ? static_cast<size_t>(ry) * w + cx : 0;
            qv[k] = gptr(qp)[p];
            if (L0) ov[k] = gptr(ch.owner)[p];
            else wv[k] = gptr(static_cast<const _Float16*>(it.wt[l]))[p];
        }
    };
    int e = next_tile(blockIdx.x);
    if (e < ntiles) fetch(e);
    while (e < ntiles) {
        const RTile tl = tiles[e];
        const RItem& it = items[tl.id];
        const RChunk& ch = chunks[it.chunk];
        const int ow = ch.lw[l + 1], oh = ch.lh[l + 1];
        const int ox0 = tl.tx * kRedTW, oy0 = tl.ty * kRedTH;
        const int local = it.local;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = threadIdx.x + 256 * k;
            if (i >= NE) continue;
            h4 z;
            z.x = z.y = z.z = z.w = static_cast<_Float16>(0.0f);
            sq[i] = inb[k] ? qv[k] : z;
            if (L0) sw[i] = static_cast<_Float16>((inb[k] && ov[k] == local) ? 1.0f : 0.0f);
            else sw[i] = inb[k] ? wv[k] : static_cast<_Float16>(0.0f);
        }
        lds_barrier();  // window staged; the previous tile's vertical pass is done with tmp
        const int nx = next_tile(e + gridDim.x);
        if (nx < ntiles) fetch(nx);
        // horizontal pass: each thread blurs two neighbouring outputs of a
        // window row (taps 2 xx .. 2 xx + 6, shared)
        for (int idx = threadIdx.x; idx < kRedRows * (kRedTW / 2); idx += 256) {
            const int rr = idx / (kRedTW / 2), xx = 2 * (idx % (kRedTW / 2));
            if (ox0 + xx >= ow) continue;
            M v[7][5];
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int j = 2 * xx + k;
                const h4 q = sq[rr * kRed0Cols + j];
                // level 0: the warped plane is premultiplied already (its
                // coverage is 0 or 1 and its RGB is 0 where uncovered)
                v[k][0] = static_cast<M>(q.x);
                v[k][1] = static_cast<M>(q.y);
                v[k][2] = static_cast<M>(q.z);
                v[k][3] = static_cast<M>(q.w);
                v[k][4] = static_cast<M>(sw[rr * kRed0Cols + j]);
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                tmp[c][rr][xx] = blur5s(v[0][c], v[1][c], v[2][c], v[3][c], v[4][c]);
                tmp[c][rr][xx + 1] = blur5s(v[2][c], v[3][c], v[4][c], v[5][c], v[6][c]);
            }
        }
        lds_barrier();  // tmp complete (and every read of sq / sw done)
        // vertical pass: each thread blurs two output rows of a column
        // (window rows 4 yy .. 4 yy + 6, shared)
        for (int idx = threadIdx.x; idx < (kRedTH / 2) * kRedTW; idx += 256) {
            const int yy = 2 * (idx / kRedTW), xx = idx % kRedTW;
            const int ox = ox0 + xx, oy = oy0 + yy;
            if (ox >= ow || oy >= oh) continue;
            const int r = 2 * yy + 2;
            M o[2][5];
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                M t[7];
#pragma unroll
                for (int k = 0; k < 7; ++k) t[k] = tmp[c][r - 2 + k][xx];
                o[0][c] = blur5s(t[0], t[1], t[2], t[3], t[4]);
                o[1][c] = blur5s(t[2], t[3], t[4], t[5], t[6]);
            }
#pragma unroll
            for (int h2 = 0; h2 < 2; ++h2) {
                if (oy + h2 >= oh) continue;
                const size_t q = static_cast<size_t>(oy + h2) * ow + ox;
                const M nd[4] = {o[h2][0], o[h2][1], o[h2][2], o[h2][3]};
                st4(it.nd[l + 1], q, nd, static_cast<_Float16*>(nullptr));
                st1<_Float16, M>(it.wt[l + 1], q, o[h2][4]);
            }
        }
        e = nx;
    }
}

// ----------------------------------------------------------- compose ---
template <typename M>
struct Coord {
    int x0, x1, y0, y1;
    M wx, wy;
};

template <typename M>
__device__ __forceinline__ Coord<M> expand_at(int x, int y, int sw, int sh, int dw, int dh) {
    const M sx = (sw > 1) ? static_cast<M>(sw - 1) / static_cast<M>(max(1, dw - 1)) : M(0);
    const M sy = (sh > 1) ? static_cast<M>(sh - 1) / static_cast<M>(max(1, dh - 1)) : M(0);
    Coord<M> c;
    const M fy = sy * static_cast<M>(y);
    c.y0 = min(static_cast<int>(fy), sh - 1);
    c.y1 = min(c.y0 + 1, sh - 1);
    c.wy = fy - static_cast<M>(c.y0);
    const M fx = sx * static_cast<M>(x);
    c.x0 = min(static_cast<int>(fx), sw - 1);
    c.x1 = min(c.x0 + 1, sw - 1);
    c.wx = fx - static_cast<M>(c.x0);
    return c;
}

// expand_at with the level's scales (sw-1)/(dw-1), (sh-1)/(dh-1) computed
// once by the caller (same values)
template <typename M>
__device__ __forceinline__ M expand_scale(int s, int d) {
    return (s > 1) ? static_cast<M>(s - 1) / static_cast<M>(max(1, d - 1)) : M(0);
}
template <typename M>
__device__ __forceinline__ Coord<M> expand_at_s(int x, int y, M sx, M sy, int sw, int sh) {
    Coord<M> c;
    const M fy = sy * static_cast<M>(y);
    c.y0 = min(static_cast<int>(fy), sh - 1);
    c.y1 = min(c.y0 + 1, sh - 1);
    c.wy = fy - static_cast<M>(c.y0);
    const M fx = sx * static_cast<M>(x);
    c.x0 = min(static_cast<int>(fx), sw - 1);
    c.x1 = min(c.x0 + 1, sw - 1);
    c.wx = fx - static_cast<M>(c.x0);
    return c;
}

// 1/d for the hole-filled Gaussians of the FP16-storage forms: the
// hardware reciprocal (1 ulp), far inside the FP16 storage error; the
// FP32 and FP64 forms divide exactly.
template <typename M, typename S>
__device__ __forceinline__ M recip(M d) {
    if constexpr (std::is_same<S, _Float16>::value) return __builtin_amdgcn_rcpf(d);
    else return M(1) / d;
}

template <typename M>
__device__ __forceinline__ M lerp2m(const Coord<M>& c, M v00, M v10, M v01, M v11) {
    const M omx = M(1) - c.wx, omy = M(1) - c.wy;
    const M top = omx * v00 + c.wx * v10;
    const M bot = omx * v01 + c.wx * v11;
    return omy * top + c.wy * bot;
}

// Hole-filled Gaussian at one level-l sample; zero outside the support box.
// The plane spans the whole level, so the load is valid outside the box
// too: it is issued with no branch (the loads of neighbouring samples are
// then in flight together) and its bits are selected away there, never
// used in arithmetic.
template <typename M, typename S>
__device__ __forceinline__ void gauss_m(const void* nd, const IBox& box, int w, int x, int y, M (&g)[3]) {
    const bool in = in_box(box, x, y);
    M q[4];
    ld4(nd, static_cast<size_t>(y) * w + x, q, static_cast<S*>(nullptr));
    const M inv = (in && q[3] > static_cast<M>(1e-6)) ? M(1) / q[3] : M(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = in ? q[k] * inv : M(0);
}

template <typename M, typename S>
__device__ __forceinline__ void expand_gauss_m(const void* nd, const IBox& box, int sw, const Coord<M>& c,
                                               M (&o)[3]) {
    M a[3], b[3], d[3], e[3];
    gauss_m<M, S>(nd, box, sw, c.x0, c.y0, a);
    gauss_m<M, S>(nd, box, sw, c.x1, c.y0, b);
    gauss_m<M, S>(nd, box, sw, c.x0, c.y1, d);
    gauss_m<M, S>(nd, box, sw, c.x1, c.y1, e);
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = lerp2m<M>(c, a[k], b[k], d[k], e[k]);
}

template <typename M>
__device__ __forceinline__ void expand_recon_m(const void* rc, int sw, const Coord<M>& c, M (&o)[3]) {
    M a[4], b[4], d[4], e[4];
    ld4(rc, static_cast<size_t>(c.y0) * sw + c.x0, a, static_cast<M*>(nullptr));
    ld4(rc, static_cast<size_t>(c.y0) * sw + c.x1, b, static_cast<M*>(nullptr));
    ld4(rc, static_cast<size_t>(c.y1) * sw + c.x0, d, static_cast<M*>(nullptr));
    ld4(rc, static_cast<size_t>(c.y1) * sw + c.x1, e, static_cast<M*>(nullptr));
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = lerp2m<M>(c, a[k], b[k], d[k], e[k]);
}

constexpr int kCompTW = 64, kCompTH = 4;

// recon_l = sum_i w_i * (G_i,l - expand(G_i,l+1)) (+ expand(recon_{l+1}) below the top)
template <int MODE>
__global__ void __launch_bounds__(256) rcompose_kernel(const RItem* __restrict__ items,
                                                       const RChunk* __restrict__ chunks,
                                                       const RTile* __restrict__ tiles, int ntiles, int l,
                                                       int phase, const int* __restrict__ cflag) {
    constexpr int MM = MODE == kModeBatched ? kModePrecision : MODE;
    using M = typename Form<MM>::M;
    using S = typename Form<MM>::S;
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RChunk& ch = chunks[tl.id];
        // compose at level l is needed below K (phase 1) and at K and above
        // for chunks that fall back (phase 2)
        if (phase == 1 && l >= (ch.sk >= 1 ? ch.sk : kMaxLevels)) continue;
        if (phase == 2 && (ch.sk < 1 || l < ch.sk || !cflag[tl.id])) continue;
        const int x = tl.tx * kCompTW + (threadIdx.x & 63), y = tl.ty * kCompTH + (threadIdx.x >> 6);
        const int w = ch.lw[l], h = ch.lh[l];
        if (x >= w || y >= h) continue;
        const bool top = (l == ch.levels - 1);
        const int nw = top ? 1 : ch.lw[l + 1], nh = top ? 1 : ch.lh[l + 1];
        Coord<M> c{};
        if (!top) c = expand_at<M>(x, y, nw, nh, w, h);
        M band[3] = {M(0), M(0), M(0)};
        const size_t p = static_cast<size_t>(y) * w + x;
        // a hole-filled Gaussian from a raw numerator/denominator sample (the
        // same arithmetic as gauss_m)
        using Raw4 = S __attribute__((ext_vector_type(4)));
        auto fill = [](const Raw4& r, bool in, M (&g)[3]) {
            M q[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) q[j] = static_cast<M>(r[j]);
            const M inv = (in && q[3] > static_cast<M>(1e-6)) ? M(1) / q[3] : M(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) g[j] = in ? q[j] * inv : M(0);
        };
        for (int k = 0; k < ch.count; ++k) {
            const RItem& it = items[ch.first + k];
            const IBox bl = box_of(it.lbox[l]);
            if (!in_box(bl, x, y)) continue;
            const M wt = ld1<M, S>(it.wt[l], p);
            if (wt == M(0)) continue;
            // the sample and (below the top) the four upsample corners, all
            // loaded raw before any is used: one memory latency, not two
            // (the planes span their whole level, so every corner address
            // is valid; a corner outside the box is selected away)
            const Raw4 qc = gptr(static_cast<const Raw4*>(it.nd[l]))[p];
            Raw4 qe[4] = {};
            if (!top) {
                const Raw4* np = static_cast<const Raw4*>(it.nd[l + 1]);
                qe[0] = gptr(np)[static_cast<size_t>(c.y0) * nw + c.x0];
                qe[1] = gptr(np)[static_cast<size_t>(c.y0) * nw + c.x1];
                qe[2] = gptr(np)[static_cast<size_t>(c.y1) * nw + c.x0];
                qe[3] = gptr(np)[static_cast<size_t>(c.y1) * nw + c.x1];
            }
            M g[3];
            fill(qc, true, g);
            if (!top) {
                const IBox bu = box_of(it.lbox[l + 1]);
                M a[3], b[3], d[3], e[3];
                fill(qe[0], in_box(bu, c.x0, c.y0), a);
                fill(qe[1], in_box(bu, c.x1, c.y0), b);
                fill(qe[2], in_box(bu, c.x0, c.y1), d);
                fill(qe[3], in_box(bu, c.x1, c.y1), e);
#pragma unroll
                for (int q = 0; q < 3; ++q) g[q] = g[q] - lerp2m<M>(c, a[q], b[q], d[q], e[q]);
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) band[q] = band[q] + wt * g[q];
        }
        if (!top) {
            M ur[3];
            expand_recon_m<M>(ch.recon[l + 1], nw, c, ur);
#pragma unroll
            for (int q = 0; q < 3; ++q) band[q] = band[q] + ur[q];
        }
        const M o4[4] = {band[0], band[1], band[2], M(0)};
        st4(ch.recon[l], p, o4, static_cast<M*>(nullptr));
    }
}

// ---------------------------------------------------------- surrogate ---
// Per level-K pixel of a chunk (tiles: 64x4 at level K), the network's
// inputs: the Gaussians of the two largest-weight contributors (ties to
// the lower index; the second is the first where only one weighs), the
// seam mask wA / (wA + wB), the single-level composite sum_i w_i G_i; the
// two Laplacian scales of the composite (channels 10-15) come from
// rsblur_kernel.
template <int MODE>
__global__ void __launch_bounds__(256) rsfeat_kernel(const RItem* __restrict__ items,
                                                     const RChunk* __restrict__ chunks,
                                                     const RTile* __restrict__ tiles, int ntiles, SurrBufs sb) {
    constexpr int MM = MODE == kModeBatched ? kModePrecision : MODE;
    using M = typename Form<MM>::M;
    using S = typename Form<MM>::S;
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RChunk& ch = chunks[tl.id];
        const int l = ch.sk;
        const int x = tl.tx * kCompTW + (threadIdx.x & 63), y = tl.ty * kCompTH + (threadIdx.x >> 6);
        const int w = ch.lw[l], h = ch.lh[l];
        if (x >= w || y >= h) continue;
        const size_t p = static_cast<size_t>(y) * w + x;
        M wa = M(-1), wb = M(-1), ga[3] = {M(0), M(0), M(0)}, gb[3] = {M(0), M(0), M(0)};
        M base[3] = {M(0), M(0), M(0)};
        for (int k = 0; k < ch.count; ++k) {
            const RItem& it = items[ch.first + k];
            const IBox bl = box_of(it.lbox[l]);
            if (!in_box(bl, x, y)) continue;
            const M wt = ld1<M, S>(it.wt[l], p);
            if (wt == M(0)) continue;
            M g[3];
            gauss_m<M, S>(it.nd[l], bl, w, x, y, g);
#pragma unroll
            for (int q = 0; q < 3; ++q) base[q] = base[q] + wt * g[q];
            if (wt > wa) {
                wb = wa;
#pragma unroll
                for (int q = 0; q < 3; ++q) { gb[q] = ga[q]; ga[q] = g[q]; }
                wa = wt;
            } else if (wt > wb) {
                wb = wt;
#pragma unroll
                for (int q = 0; q < 3; ++q) gb[q] = g[q];
            }
        }
        if (wb < M(0)) {  // one contributor (or none)
#pragma unroll
            for (int q = 0; q < 3; ++q) gb[q] = ga[q];
            wb = M(0);
        }
        if (wa < M(0)) wa = M(0);
        const M m = (wa + wb > M(0)) ? wa / (wa + wb) : M(0);
        float* f = sb.feat + (static_cast<size_t>(ch.spix0) + p) * kSurrIn;
        const float v[16] = {static_cast<float>(ga[0]), static_cast<float>(ga[1]), static_cast<float>(ga[2]),
                             static_cast<float>(gb[0]), static_cast<float>(gb[1]), static_cast<float>(gb[2]),
                             static_cast<float>(m), static_cast<float>(base[0]), static_cast<float>(base[1]),
                             static_cast<float>(base[2]), 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4)
            *reinterpret_cast<float4*>(f + 4 * q4) = make_float4(v[4 * q4], v[4 * q4 + 1], v[4 * q4 + 2], v[4 * q4 + 3]);
    }
}

// [1,4,6,4,1]/16 (x) [1,4,6,4,1]/16 blur of a 3-channel level-K plane of a
// chunk at dilation `dil`, replicate borders.
__device__ __forceinline__ void blur25(const float* src, int stride, int ch0, int w, int h, int x, int y, int dil,
                                       float (&o)[3]) {
    const float k[5] = {1.f / 16, 4.f / 16, 6.f / 16, 4.f / 16, 1.f / 16};
    o[0] = o[1] = o[2] = 0.f;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int yy = min(max(y + (j - 2) * dil, 0), h - 1);
        float r[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const int xx = min(max(x + (i - 2) * dil, 0), w - 1);
            const float* q = src + (static_cast<size_t>(yy) * w + xx) * stride + ch0;
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] = fmaf(k[i], q[c], r[c]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = fmaf(k[j], r[c], o[c]);
    }
}

// pass 0: aux = B1 = blur(composite);  pass 1: feat[10..12] = composite - B1,
// feat[13..15] = B1 - blur_dilated2(B1).
__global__ void __launch_bounds__(256) rsblur_kernel(const RChunk* __restrict__ chunks,
                                                     const RTile* __restrict__ tiles, int ntiles, SurrBufs sb,
                                                     int pass) {
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RChunk& ch = chunks[tl.id];
        const int l = ch.sk;
        const int x = tl.tx * kCompTW + (threadIdx.x & 63), y = tl.ty * kCompTH + (threadIdx.x >> 6);
        const int w = ch.lw[l], h = ch.lh[l];
        if (x >= w || y >= h) continue;
        const size_t q = static_cast<size_t>(ch.spix0) + static_cast<size_t>(y) * w + x;
        const float* feat = sb.feat + static_cast<size_t>(ch.spix0) * kSurrIn;
        const float* aux = sb.aux + static_cast<size_t>(ch.spix0) * 4;
        float o[3];
        if (pass == 0) {
            blur25(feat, kSurrIn, 7, w, h, x, y, 1, o);
            *reinterpret_cast<float4*>(sb.aux + q * 4) = make_float4(o[0], o[1], o[2], 0.f);
        } else {
            blur25(aux, 4, 0, w, h, x, y, 2, o);
            const float4 b1 = *reinterpret_cast<const float4*>(sb.aux + q * 4);
            float* f = sb.feat + q * kSurrIn;
            const float4 c = *reinterpret_cast<const float4*>(f + 4);  // 4..7
            const float4 d = *reinterpret_cast<const float4*>(f + 8);  // 8..11
            const float base[3] = {c.w, d.x, d.y};
            *reinterpret_cast<float4*>(f + 8) = make_float4(d.x, d.y, base[0] - b1.x, base[1] - b1.y);
            *reinterpret_cast<float4*>(f + 12) = make_float4(base[2] - b1.z, b1.x - o[0], b1.y - o[1], b1.z - o[2]);
        }
    }
}

// Chunk flags start at 0 for chunks with a surrogate level and at 1 (plain
// blend throughout) for chunks without one.
__global__ void rscflag_init_kernel(const RChunk* __restrict__ chunks, int nch, int* __restrict__ cflag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nch) cflag[c] = chunks[c].sk >= 1 ? 0 : 1;
}

// One block per 16x16 surrogate tile: fall back when the mean predicted
// squared error of the low band (10^out[3]) exceeds `thr`, or (thrMax > 0)
// when any pixel's exceeds `thrMax`.
__global__ void __launch_bounds__(256) rsflag_kernel(const SurrImg* __restrict__ imgs,
                                                     const SurrTile* __restrict__ tiles, int t0, SurrBufs sb,
                                                     float thr, float thrMax) {
    __shared__ float ssum[256];
    __shared__ float smax[256];
    __shared__ int scnt[256];
    const SurrTile tl = tiles[t0 + blockIdx.x];
    const SurrImg im = imgs[tl.img];
    const int x = tl.tx * 16 + (threadIdx.x & 15), y = tl.ty * 16 + (threadIdx.x >> 4);
    float e = 0.f;
    int n = 0;
    if (x < im.w && y < im.h) {
        const size_t q = static_cast<size_t>(im.pix0) + static_cast<size_t>(y) * im.w + x;
        e = exp10f(fminf(sb.out[q * 4 + 3], 10.f));
        n = 1;
    }
    ssum[threadIdx.x] = e;
    smax[threadIdx.x] = e;
    scnt[threadIdx.x] = n;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            ssum[threadIdx.x] += ssum[threadIdx.x + k];
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + k]);
            scnt[threadIdx.x] += scnt[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool overMean = ssum[0] > thr * static_cast<float>(max(scnt[0], 1));
        const bool overMax = thrMax > 0.f && smax[0] > thrMax;
        const int f = (overMean || overMax) ? 1 : 0;
        sb.tflag[t0 + blockIdx.x] = f;
        if (f) atomicOr(&sb.cflag[im.tag], 1);
    }
}

// recon_K = composite + residual on every tile that does not fall back.
template <int MODE>
__global__ void __launch_bounds__(256) rsapply_kernel(const RChunk* __restrict__ chunks,
                                                      const RTile* __restrict__ tiles, int ntiles, SurrBufs sb) {
    constexpr int MM = MODE == kModeBatched ? kModePrecision : MODE;
    using M = typename Form<MM>::M;
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RChunk& ch = chunks[tl.id];
        const int l = ch.sk;
        const int x = tl.tx * kCompTW + (threadIdx.x & 63), y = tl.ty * kCompTH + (threadIdx.x >> 6);
        const int w = ch.lw[l], h = ch.lh[l];
        if (x >= w || y >= h) continue;
        if (sb.cflag[tl.id] && sb.tflag[ch.stile0 + (y >> 4) * ch.stx + (x >> 4)]) continue;
        const size_t p = static_cast<size_t>(y) * w + x;
        const size_t q = static_cast<size_t>(ch.spix0) + p;
        const float* f = sb.feat + q * kSurrIn;
        const float4 o = *reinterpret_cast<const float4*>(sb.out + q * 4);
        const M v[4] = {f[7] + kSurrResScale * o.x, f[8] + kSurrResScale * o.y, f[9] + kSurrResScale * o.z, M(0)};
        st4(ch.recon[l], p, v, static_cast<M*>(nullptr));
    }
}

// Training-data export: the plain blend's level-K reconstruction.
template <int MODE>
__global__ void __launch_bounds__(256) rsexport_kernel(const RChunk* __restrict__ chunks,
                                                       const RTile* __restrict__ tiles, int ntiles, SurrBufs sb) {
    constexpr int MM = MODE == kModeBatched ? kModePrecision : MODE;
    using M = typename Form<MM>::M;
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RChunk& ch = chunks[tl.id];
        const int l = ch.sk;
        const int x = tl.tx * kCompTW + (threadIdx.x & 63), y = tl.ty * kCompTH + (threadIdx.x >> 6);
        const int w = ch.lw[l], h = ch.lh[l];
        if (x >= w || y >= h) continue;
        const size_t p = static_cast<size_t>(y) * w + x;
        M r[4];
        ld4(ch.recon[l], p, r, static_cast<M*>(nullptr));
        float* t = sb.target + (static_cast<size_t>(ch.spix0) + p) * 3;
        t[0] = r[0]; t[1] = r[1]; t[2] = r[2];
    }
}

// Level 0: only the owner carries weight (0/1 owner mask); FP32 RGB out.
// A 256-thread block takes one 64x16 tile per iteration, each thread four
// pixels of one column (rows r, r + 4, r + 8, r + 12):
//   A. the chunk's contributor fields go to LDS; each thread loads its
//      owners and ORs them into the tile's owner mask;
//   B. the level-1 region the tile interpolates from (at most 34 x 10
//      texels: the level-1 scale is <= 1/2) is staged in LDS once: the
//      reconstruction, and the hole-filled Gaussians of the first
//      kFinSlots owners of the tile (a contributor's slot is its rank in
//      the mask);
//   C. every pixel interpolates from LDS (a pixel whose owner has no slot
//      gathers that owner's Gaussians from memory).
// Each texel's Gaussian is formed once instead of once per pixel; the
// values are the ones the per-pixel evaluation gives.
struct FinalItem {
    const void* warped;
    const void* nd1;
    int box[4];
};
constexpr int kFinSlots = 3;  // (3 slots of RGB keep the block's LDS at ~22 KB: 6 blocks per CU)
constexpr int kFinW1 = 36, kFinH1 = 12;  // staged region bound (34 x 10 needed)

template <int MODE>
__global__ void __launch_bounds__(256) rfinal_kernel(const RItem* __restrict__ items,
                                                     const RChunk* __restrict__ chunks,
                                                     const RTile* __restrict__ tiles, int ntiles) {
    constexpr int MM = MODE == kModeBatched ? kModePrecision : MODE;
    using M = typename Form<MM>::M;
    using S = typename Form<MM>::S;
    constexpr int K = 4;
    constexpr int kRowStep = kWarpTH / K;
    __shared__ FinalItem sit[I360HIP_MAX_CONTRIBUTORS];
    // per slot: the level-1 difference R1 - G1 of the slot's contributor
    // (a fourth, unused component aligns an element for one 16-byte access)
    __shared__ M sG[kFinSlots][kFinH1 * kFinW1][4];
    __shared__ uint32_t smask;
    const int tid = static_cast<int>(threadIdx.x);
    const int lx = tid & 63, ry = tid >> 6;
    if (tid == 0) smask = 0u;
    lds_barrier();  // the mask is clear before any thread ORs into it
    for (int e = blockIdx.x; e < ntiles; e += gridDim.x) {
        const RTile tl = tiles[e];
        const RChunk& ch = chunks[tl.id];
        const int w = ch.w, h = ch.h, count = ch.count;
        // A
        if (tid < count) {
            const RItem& it = items[ch.first + tid];
            FinalItem f;
            f.warped = it.warped;
            f.nd1 = it.nd[1];
            f.box[0] = it.lbox[1][0]; f.box[1] = it.lbox[1][1]; f.box[2] = it.lbox[1][2]; f.box[3] = it.lbox[1][3];
            sit[tid] = f;
        }
        const int x = tl.tx * kWarpTW + lx;
        bool ok[K];
        size_t p[K];
        int o[K];
        uint32_t bits = 0u;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int y = tl.ty * kWarpTH + ry + kRowStep * k;
            ok[k] = x < w && y < h;
            p[k] = ok[k] ? static_cast<size_t>(y) * w + x : 0;
            const int v = gptr(ch.owner)[p[k]];
            o[k] = ok[k] ? v : -1;
            bits |= o[k] >= 0 ? (1u << o[k]) : 0u;
        }
        if (bits) atomicOr(&smask, bits);
        lds_barrier();
        const uint32_t mask = smask;
        // the warped values, in flight across the level-1 staging: loaded
        // raw (storage type) and converted where they are used, so that no
        // instruction needs them, and none of the four loads waits, before
        // the blend
        using Raw4 = S __attribute__((ext_vector_type(4)));
        Raw4 qr[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const FinalItem& it = sit[o[k] >= 0 ? o[k] : 0];
            qr[k] = Raw4{};
            if (count > 0) qr[k] = gptr(static_cast<const Raw4*>(it.warped))[p[k]];
        }
        const bool deep = ch.levels > 1;  // tile-uniform
        const int lw1 = deep ? ch.lw[1] : 1, lh1 = deep ? ch.lh[1] : 1;
        const M sx = expand_scale<M>(lw1, w), sy = expand_scale<M>(lh1, h);
        // the tile's level-1 region
        const Coord<M> cA = expand_at_s<M>(tl.tx * kWarpTW, tl.ty * kWarpTH, sx, sy, lw1, lh1);
        const Coord<M> cB = expand_at_s<M>(min(tl.tx * kWarpTW + kWarpTW - 1, w - 1),
                                           min(tl.ty * kWarpTH + kWarpTH - 1, h - 1), sx, sy, lw1, lh1);
        const int X0 = cA.x0, Y0 = cA.y0, W1 = cB.x1 - cA.x0 + 1, H1 = cB.y1 - cA.y0 + 1;
        const bool staged = deep && count > 0 && W1 <= kFinW1 && H1 <= kFinH1;  // tile-uniform
        const int nslots = min(__popc(mask), kFinSlots);
        // B: per slot, the level-1 difference R1 - G1 of the slot's
        // contributor over the tile's level-1 region, in a fixed
        // kFinH1 x kFinW1 layout (constant divisors; elements beyond
        // W1 x H1 are skipped).  A slot's loads of a thread are issued
        // before any is used, so they are in flight together.
        if (staged) {
            constexpr int kSlotN = kFinH1 * kFinW1;
            constexpr int kIt = (kSlotN + 255) / 256;
            // Loads are raw (storage type) and unconditional: an element
            // outside the region reads texel 0 of the level-1 plane and is
            // skipped at the write.  The reconstruction is the chunk's, the
            // same for every slot, so it is loaded once.
            using RawS = S __attribute__((ext_vector_type(4)));
            using RawM = M __attribute__((ext_vector_type(4)));
            int qq[kIt], tx1[kIt], ty1[kIt];
            size_t t1[kIt];
            bool use[kIt];
            RawM rr[kIt];
#pragma unroll
            for (int k = 0; k < kIt; ++k) {
                qq[k] = tid + 256 * k;
                const int qy = qq[k] / kFinW1, qx = qq[k] - qy * kFinW1;
                use[k] = qq[k] < kSlotN && qy < H1 && qx < W1;
                ty1[k] = Y0 + qy;
                tx1[k] = X0 + qx;
                t1[k] = use[k] ? static_cast<size_t>(ty1[k]) * lw1 + tx1[k] : 0;
                rr[k] = gptr(static_cast<const RawM*>(ch.recon[1]))[t1[k]];
            }
            uint32_t m = mask;
            for (int sl = 0; sl < nslots; ++sl, m &= m - 1u) {
                const FinalItem& it = sit[__ffs(m) - 1];
                RawS ar[kIt];
#pragma unroll
                for (int k = 0; k < kIt; ++k) ar[k] = gptr(static_cast<const RawS*>(it.nd1))[t1[k]];
#pragma unroll
                for (int k = 0; k < kIt; ++k) {
                    if (!use[k]) continue;
                    M a[4], r[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        a[j] = static_cast<M>(ar[k][j]);
                        r[j] = rr[k][j];
                    }
                    const bool in = tx1[k] >= it.box[0] && tx1[k] <= it.box[2] && ty1[k] >= it.box[1] &&
                                    ty1[k] <= it.box[3];
                    const M inv = (a[3] > static_cast<M>(1e-6)) ? recip<M, S>(a[3]) : M(0);
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) sG[sl][qq[k]][c3] = r[c3] - (in ? a[c3] * inv : M(0));
                    sG[sl][qq[k]][3] = M(0);
                }
            }
        }
        lds_barrier();
        // C
        if (ch.owner_out) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (ok[k]) gptr(ch.owner_out)[p[k]] = o[k];
        }
        // every owner of the tile has a staged slot (tile-uniform): no
        // per-pixel branch; an unowned pixel's value is computed from slot 0
        // and discarded (its reads stay in the staged region: the tile's
        // coordinates, clamped to the level-1 grid, lie inside it)
        const bool fast = staged && __popc(mask) <= kFinSlots;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            M v[3] = {M(0), M(0), M(0)};
            if (fast) {
                const int y = tl.ty * kWarpTH + ry + kRowStep * k;
                const Coord<M> c = expand_at_s<M>(x, y, sx, sy, lw1, lh1);
                const int slot = __popc(mask & ((1u << max(o[k], 0)) - 1u));
                // small non-negative operands: 24-bit multiplies (full rate)
                const int r0 = __mul24(slot, kFinH1 * kFinW1) + __mul24(c.y0 - Y0, kFinW1) - X0;
                const int r1 = r0 + __mul24(c.y1 - c.y0, kFinW1);
                const int qs[4] = {r0 + c.x0, r0 + c.x1, r1 + c.x0, r1 + c.x1};
                M dq[4][4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int c3 = 0; c3 < 4; ++c3) dq[j][c3] = sG[0][qs[j]][c3];
#pragma unroll
                for (int c3 = 0; c3 < 3; ++c3)
                    v[c3] = static_cast<M>(qr[k][c3]) + lerp2m<M>(c, dq[0][c3], dq[1][c3], dq[2][c3], dq[3][c3]);
            } else if (o[k] >= 0) {
                // the owner covers the pixel: its coverage (alpha) is exactly 1
                // and its hole-filled level-0 Gaussian is the warped value
#pragma unroll
                for (int c3 = 0; c3 < 3; ++c3) v[c3] = static_cast<M>(qr[k][c3]);
                if (deep) {
                    const int y = tl.ty * kWarpTH + ry + kRowStep * k;
                    const Coord<M> c = expand_at_s<M>(x, y, sx, sy, lw1, lh1);
                    const int slot = __popc(mask & ((1u << o[k]) - 1u));
                    // level 0 = G0 - Up(G1) + Up(R1) = G0 + Up(R1 - G1): the
                    // bilinear upsample is linear, so one interpolation of
                    // the level-1 difference (equal up to rounding)
                    M dq[4][3];
                    const int xs[4] = {c.x0, c.x1, c.x0, c.x1};
                    const int ys[4] = {c.y0, c.y0, c.y1, c.y1};
                    if (staged && slot < kFinSlots) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int q = (ys[j] - Y0) * kFinW1 + (xs[j] - X0);
#pragma unroll
                            for (int c3 = 0; c3 < 3; ++c3) dq[j][c3] = sG[slot][q][c3];
                        }
                    } else {
                        const FinalItem& it = sit[o[k]];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            M a[4], rr[4];
                            const size_t t = static_cast<size_t>(ys[j]) * lw1 + xs[j];
                            ld4(it.nd1, t, a, static_cast<S*>(nullptr));
                            ld4(ch.recon[1], t, rr, static_cast<M*>(nullptr));
                            const bool in = xs[j] >= it.box[0] && xs[j] <= it.box[2] && ys[j] >= it.box[1] && ys[j] <= it.box[3];
                            const M inv = (a[3] > static_cast<M>(1e-6)) ? recip<M, S>(a[3]) : M(0);
#pragma unroll
                            for (int c3 = 0; c3 < 3; ++c3) dq[j][c3] = rr[c3] - (in ? a[c3] * inv : M(0));
                        }
                    }
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) v[c3] = v[c3] + lerp2m<M>(c, dq[0][c3], dq[1][c3], dq[2][c3], dq[3][c3]);
                }
            }
            if (ok[k]) {
                float f[3];
#pragma unroll
                for (int c3 = 0; c3 < 3; ++c3) {
                    M t = v[c3] < M(1) ? v[c3] : M(1);
                    t = M(0) < t ? t : M(0);
                    f[c3] = o[k] < 0 ? 0.0f : static_cast<float>(t);
                }
                if (ch.out16) {
                    // RGB48 payload: u16(clamp(v * 65535) + 0.5) of the FP32 value (SeamBlend.cpp:529-534)
                    auto* d16 = gptr(reinterpret_cast<uint16_t*>(ch.out) + p[k] * 3);
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) {
                        const float w = f[c3] * 65535.0f;
                        const float lo = (w < 65535.0f) ? w : 65535.0f;
                        d16[c3] = static_cast<uint16_t>(static_cast<unsigned int>(((0.0f < lo) ? lo : 0.0f) + 0.5f));
                    }
                } else {
                    auto* dst = gptr(ch.out + p[k] * 3);
#pragma unroll
                    for (int c3 = 0; c3 < 3; ++c3) dst[c3] = f[c3];
                }
            }
        }
        if (tid == 0) smask = 0u;  // read by every thread before the barrier above
        lds_barrier();             // the next tile reuses sit, sR, sG and the mask
    }
}

// --------------------------------------------------------------- plan ---
// One seam pair of a chunk: contributors i < j (chunk-local), its round
// and its scratch slot within that round.
struct SeamPair {
    int i, j, round, slot;
};

enum SurrMode : int { kSurrNone = 0, kSurrRender = 1, kSurrExport = 2 };

// What a surrogate render (or export) adds to a batched render.
struct SurrRun {
    int mode = kSurrNone;
    SurrBufs bufs{};
    const uint16_t* weights = nullptr;
    const float* biases = nullptr;
    float threshold = 0.0f;  // mean predicted squared error above which a tile falls back
    float max_threshold = 0.0f;  // per-pixel predicted squared error above which it falls back (0: off)
};

struct Plan {
    std::vector<RItem> items;
    std::vector<RChunk> chunks;
    std::vector<ChamferWork> chamfer;
    // tiles per stage, grouped by chunk: [chunk] -> (offset, count)
    std::vector<RTile> tiles;
    struct Range { int off = 0, n = 0; };
    std::vector<Range> warp, own, fin;
    std::vector<std::vector<Range>> red;   // [chunk][l] tiles of level l+1
    std::vector<std::vector<Range>> comp;  // [chunk][l] tiles of level l
    std::vector<std::vector<SeamPair>> pairs;  // [chunk] item-local pairs, reference order
    size_t bytes = 0;
    int max_w = 0;
    int mode = 0;
    uint64_t key = 0;
    // surrogate (batched form only)
    int surr = kSurrNone;
    int snpix = 0;                   // level-K pixels over all chunks
    std::vector<SurrImg> simgs;      // one per chunk with a surrogate level
    std::vector<SurrTile> stiles;    // 16x16 tiles of those images
    Range sfeat;                     // 64x4 level-K tiles (RTile, id = chunk)
    std::vector<Range> sfeat_c;      // ... of each chunk (in chunk order, within sfeat)
    std::vector<Range> stile_c;      // each chunk's 16x16 surrogate tiles (in chunk order)
    SurrBufs sbufs{};                // offsets until patched
};

inline size_t A256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Groups a chunk's seam pairs into rounds.  A pair (i, j) reads only which
// of i / j owns a pixel (owners outside {i, j} count as "neither") and
// rewrites only pixels owned by i or j, to i or j (SeamBlend.cpp:336-343,
// 437-442), and it reads and rewrites owners only inside the overlap of
// i's and j's coverage boxes (SeamBlend.cpp:313-325, 429-442); its bands,
// costs and sides come from coverage, images and boxes, which no pair
// changes.  Two pairs commute, then, when they have no contributor in
// common, or when the overlaps of their support boxes (which hold the
// coverage boxes) are disjoint.  Any order that keeps the reference order
// among the other pairs gives the reference's owner map: a pair's round is
// one past the latest round of an earlier pair it does not commute with,
// and the pairs of one round run side by side, each in its own scratch
// slot.  `boxes` holds each contributor's support box.  Returns the
// number of slots (largest round).
int schedule_pairs(std::vector<SeamPair>& pr, const std::vector<std::array<int, 4>>& boxes) {
    std::vector<int> fill;
    int slots = 0;
    auto overlap = [&](const SeamPair& p, int (&o)[4]) {
        const auto& a = boxes[p.i];
        const auto& b = boxes[p.j];
        o[0] = std::max(a[0], b[0]); o[1] = std::max(a[1], b[1]);
        o[2] = std::min(a[2], b[2]); o[3] = std::min(a[3], b[3]);
    };
    for (size_t k = 0; k < pr.size(); ++k) {
        SeamPair& p = pr[k];
        int op[4];
        overlap(p, op);
        int r = 0;
        for (size_t q = 0; q < k; ++q) {
            const SeamPair& e = pr[q];
            if (e.i != p.i && e.i != p.j && e.j != p.i && e.j != p.j) continue;  // no contributor in common
            int oe[4];
            overlap(e, oe);
            if (std::max(op[0], oe[0]) > std::min(op[2], oe[2]) || std::max(op[1], oe[1]) > std::min(op[3], oe[3]))
                continue;  // disjoint overlaps
            r = std::max(r, e.round + 1);
        }
        p.round = r;
        if (static_cast<int>(fill.size()) <= r) fill.resize(r + 1, 0);
        p.slot = fill[r]++;
        slots = std::max(slots, fill[r]);
    }
    return slots;
}

void dilate(const int (&in)[4], int ow, int oh, int (&out)[4]) {
    // support of [2x-2, 2x+2] intersecting [in0, in1]
    auto lo = [](int a) { return a - 2 >= 0 ? (a - 2 + 1) / 2 : 0; };
    out[0] = std::max(0, lo(in[0]));
    out[1] = std::max(0, lo(in[1]));
    out[2] = std::min(ow - 1, (in[2] + 2) / 2);
    out[3] = std::min(oh - 1, (in[3] + 2) / 2);
}

bool invert_f32(const double* hd, float* inv) {
    float hm[9];
    for (int i = 0; i < 9; ++i) hm[i] = static_cast<float>(hd[i]);
    const float a = hm[0], b = hm[1], c = hm[2], d = hm[3], e = hm[4], f = hm[5], g = hm[6], k = hm[7], j = hm[8];
    const float c0 = e * j - f * k, c1 = d * j - f * g, c2 = d * k - e * g;
    const float det = (a * c0 - b * c1) + c * c2;
    if (std::fabs(det) < 1e-12f) return false;
    const float s = 1.0f / det;
    inv[0] = c0 * s; inv[1] = (c * k - b * j) * s; inv[2] = (b * f - c * e) * s;
    inv[3] = (f * g - d * j) * s; inv[4] = (a * j - c * g) * s; inv[5] = (c * d - a * f) * s;
    inv[6] = c2 * s; inv[7] = (b * g - a * k) * s; inv[8] = (a * e - b * d) * s;
    return true;
}

uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}

}  // namespace

// Builds the plan (host tables + workspace layout); returns false on error.
// `caller`: the caller's index of each planned chunk, for messages (null:
// the planned order is the caller's).
static bool build_plan(const i360hip_frame* frames, int nf, const i360hip_chunk* chunks, int nc, const int* contrib,
                       int bands, int mode, Plan& P, Status& st, int surr = kSurrNone, bool out16 = false,
                       const int* caller = nullptr) {
    P = Plan{};
    P.mode = mode;
    P.surr = (mode == kModeBatched) ? surr : kSurrNone;
    const bool batched = mode == kModeBatched;
    const size_t sS = (mode == kModeFP64) ? 8 : (mode == kModePlain ? 4 : 2);
    const size_t sM = (mode == kModeFP64) ? 8 : 4;
    uint64_t key = mix(0, static_cast<uint64_t>(mode));
    key = mix(key, static_cast<uint64_t>(bands));
    key = mix(key, static_cast<uint64_t>(P.surr));
    key = mix(key, out16 ? 1u : 0u);
    // layout offsets are relative; absolute pointers are patched after allocation
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += A256(b); return o; };
    struct Pending { size_t warped, scov, dist, nd[kMaxLevels], wt[kMaxLevels]; };
    struct PendingC { size_t boxes, owner, recon[kMaxLevels], lohi, prevl, chosen, back, costs, counters; };
    std::vector<Pending> pit;
    std::vector<PendingC> pch;
    // the per-chunk forms reuse one chunk's worth of buffers
    size_t slot_base = 0, slot_max = 0;
    for (int ci = 0; ci < nc; ++ci) {
        const i360hip_chunk& cc = chunks[ci];
        if (cc.width <= 0 || cc.height <= 0 || !cc.out || cc.count < 0 || cc.first < 0) {
            st.set(I360HIP_INVALID_PARAM, "chunk %d: invalid rectangle, output or contributor range", caller ? caller[ci] : ci);
            return false;
        }
        if (cc.width > 16384) {
            st.set(I360HIP_INVALID_PARAM, "chunk %d wider than 16384 px", caller ? caller[ci] : ci);
            return false;
        }
        if (static_cast<uint64_t>(cc.width) * static_cast<uint64_t>(cc.height) >= (1ull << 31)) {
            st.set(I360HIP_INVALID_PARAM, "chunk %d has 2^31 pixels or more", caller ? caller[ci] : ci);
            return false;
        }
        key = mix(key, static_cast<uint64_t>(cc.x) * 1000003ull + cc.y);
        key = mix(key, static_cast<uint64_t>(cc.width) * 1000003ull + cc.height);
        key = mix(key, reinterpret_cast<uintptr_t>(cc.out));
        key = mix(key, reinterpret_cast<uintptr_t>(cc.owner));
        if (!batched) off = slot_base;
        RChunk ch{};
        ch.x = cc.x; ch.y = cc.y; ch.w = cc.width; ch.h = cc.height;
        ch.out = cc.out;
        ch.out16 = out16 ? 1 : 0;
        ch.owner_out = cc.owner;
        ch.levels = ClampBands(cc.width, cc.height, bands);
        ch.lw[0] = cc.width; ch.lh[0] = cc.height;
        for (int l = 1; l < ch.levels; ++l) { ch.lw[l] = (ch.lw[l - 1] + 1) / 2; ch.lh[l] = (ch.lh[l - 1] + 1) / 2; }
        ch.sk = (P.surr != kSurrNone && ch.levels >= 4) ? ch.levels - 3 : 0;
        ch.first = static_cast<int>(P.items.size());
        const size_t n = static_cast<size_t>(cc.width) * cc.height;
        std::vector<std::pair<int, int>> used;  // (frame, local)
        const int cap = std::min(cc.count, I360HIP_MAX_CONTRIBUTORS);
        for (int k = 0; k < cap; ++k) {
            const int f = contrib[cc.first + k];
            key = mix(key, static_cast<uint64_t>(f));
            if (f < 0 || f >= nf) {
                st.set(I360HIP_INVALID_PARAM, "chunk %d: contributor %d out of range", caller ? caller[ci] : ci, f);
                return false;
            }
            const i360hip_frame& fr = frames[f];
            if (!fr.pixels || fr.width < 2 || fr.height < 2 ||
                (fr.format != I360HIP_PIX_F16 && fr.format != I360HIP_PIX_F32 && fr.format != I360HIP_PIX_U16)) {
                st.set(I360HIP_INVALID_PARAM, "frame %d: invalid pixels/size/format", f);
                return false;
            }
            RItem it{};
            it.src = fr.pixels; it.fw = fr.width; it.fh = fr.height; it.fmt = fr.format; it.gain = fr.gain;
            for (int q = 0; q < 9; ++q) { it.hfwd[q] = fr.h[q]; key = mix(key, *reinterpret_cast<const uint64_t*>(&fr.h[q])); }
            key = mix(key, reinterpret_cast<uintptr_t>(fr.pixels));
            key = mix(key, static_cast<uint64_t>(fr.width) * 65537ull + static_cast<uint64_t>(fr.height) * 3ull + fr.format);
            {
                uint32_t gb;
                std::memcpy(&gb, &fr.gain, 4);
                key = mix(key, gb);
            }
            // footprint: AABB of the projected frame corners, widened by 2 px
            double mnx = 1e300, mny = 1e300, mxx = -1e300, mxy = -1e300;
            bool finite = true;
            for (int cnr = 0; cnr < 4; ++cnr) {
                const double u = (cnr & 1) ? fr.width : 0.0, v = (cnr & 2) ? fr.height : 0.0;
                const double q = fr.h[6] * u + fr.h[7] * v + fr.h[8];
                if (std::fabs(q) < 1e-12) { finite = false; break; }
                const double X = (fr.h[0] * u + fr.h[1] * v + fr.h[2]) / q;
                const double Y = (fr.h[3] * u + fr.h[4] * v + fr.h[5]) / q;
                mnx = std::min(mnx, X); mny = std::min(mny, Y); mxx = std::max(mxx, X); mxy = std::max(mxy, Y);
            }
            int fb[4];
            if (finite) {
                fb[0] = static_cast<int>(std::max(-1.0, std::floor(mnx) - 2 - cc.x));
                fb[1] = static_cast<int>(std::max(-1.0, std::floor(mny) - 2 - cc.y));
                fb[2] = static_cast<int>(std::min(1e9, std::ceil(mxx) + 2 - cc.x));
                fb[3] = static_cast<int>(std::min(1e9, std::ceil(mxy) + 2 - cc.y));
                fb[0] = std::max(fb[0], 0); fb[1] = std::max(fb[1], 0);
                fb[2] = std::min(fb[2], cc.width - 1); fb[3] = std::min(fb[3], cc.height - 1);
            } else {  // a corner at infinity: the whole chunk may be covered
                fb[0] = 0; fb[1] = 0; fb[2] = cc.width - 1; fb[3] = cc.height - 1;
            }
            if (fb[0] > fb[2] || fb[1] > fb[3]) continue;  // never covers this chunk
            if (mode == kModePlain && !invert_f32(fr.h, it.hinv32)) continue;  // RenderExecutor.cpp:102
            std::memcpy(it.lbox[0], fb, sizeof(fb));
            for (int l = 0; l + 1 < ch.levels; ++l) dilate(it.lbox[l], ch.lw[l + 1], ch.lh[l + 1], it.lbox[l + 1]);
            it.chunk = ci;
            it.local = static_cast<int>(used.size());
            Pending pd{};
            pd.warped = take(n * 4 * sS);
            const size_t sheared = static_cast<size_t>(DistPitch(cc.width)) * ShearRows(cc.width, cc.height);
            pd.scov = take(sheared);
            pd.dist = take(sheared * 4);
            for (int l = 1; l < ch.levels; ++l) {
                const size_t m = static_cast<size_t>(ch.lw[l]) * ch.lh[l];
                pd.nd[l] = take(m * 4 * sS);
                pd.wt[l] = take(m * sS);
            }
            pit.push_back(pd);
            P.items.push_back(it);
            used.push_back({f, it.local});
        }
        ch.count = static_cast<int>(used.size());
        PendingC pc{};
        pc.boxes = take(16 * static_cast<size_t>(std::max(1, ch.count)));
        pc.owner = take(n);
        for (int l = 1; l < ch.levels; ++l) pc.recon[l] = take(static_cast<size_t>(ch.lw[l]) * ch.lh[l] * 4 * sM);
        // pairs in the reference's order (SeamBlend.cpp:573-584); pairs whose
        // footprints are disjoint have disjoint coverage boxes, which the
        // reference skips, so they are left out here
        std::vector<SeamPair> pr;
        for (int i = 0; i < ch.count; ++i)
            for (int j = i + 1; j < ch.count; ++j) {
                const int* a = P.items[ch.first + i].lbox[0];
                const int* b = P.items[ch.first + j].lbox[0];
                if (std::max(a[0], b[0]) <= std::min(a[2], b[2]) && std::max(a[1], b[1]) <= std::min(a[3], b[3]))
                    pr.push_back({i, j, 0, 0});
            }
        std::vector<std::array<int, 4>> sboxes(ch.count);
        for (int i = 0; i < ch.count; ++i) {
            const int* a = P.items[ch.first + i].lbox[0];
            sboxes[i] = {a[0], a[1], a[2], a[3]};
        }
        const int nslots = schedule_pairs(pr, sboxes);
        const size_t lines = static_cast<size_t>(std::max(cc.width, cc.height)) * std::max(1, nslots);
        pc.lohi = take(8 * lines); pc.prevl = take(4 * lines); pc.chosen = take(4 * lines); pc.back = take(64 * lines);
        pc.costs = take(64 * 8 * lines);
        pc.counters = take(4 * static_cast<size_t>(std::max(1, ch.count)) * std::max(1, ch.count));
        pch.push_back(pc);
        P.chunks.push_back(ch);
        P.pairs.push_back(pr);
        P.max_w = std::max(P.max_w, cc.width);
        if (!batched) slot_max = std::max(slot_max, off);
    }
    if (P.surr != kSurrNone) {
        // level-K images of every chunk with a surrogate level, side by side
        P.stile_c.assign(P.chunks.size(), Plan::Range{0, 0});
        for (size_t c = 0; c < P.chunks.size(); ++c) {
            RChunk& ch = P.chunks[c];
            P.stile_c[c].off = static_cast<int>(P.stiles.size());
            if (ch.sk < 1) continue;
            const int w = ch.lw[ch.sk], h = ch.lh[ch.sk];
            ch.spix0 = P.snpix;
            ch.stile0 = static_cast<int>(P.stiles.size());
            ch.stx = (w + 15) / 16;
            const int img = static_cast<int>(P.simgs.size());
            P.simgs.push_back(SurrImg{P.snpix, w, h, static_cast<int>(c)});
            for (int ty = 0; ty < (h + 15) / 16; ++ty)
                for (int tx = 0; tx < ch.stx; ++tx) P.stiles.push_back(SurrTile{img, tx, ty, 0});
            P.stile_c[c].n = static_cast<int>(P.stiles.size()) - P.stile_c[c].off;
            P.snpix += w * h;
        }
        const size_t px = static_cast<size_t>(std::max(1, P.snpix));
        auto mk = [&](size_t b) { return reinterpret_cast<float*>(take(b)); };
        P.sbufs.feat = mk(px * kSurrIn * 4);
        P.sbufs.aux = mk(px * 16);
        P.sbufs.out = mk(px * 16);
        P.sbufs.tmp0 = mk(px * kSurrHidden * 4);
        P.sbufs.tmp1 = mk(px * kSurrHidden * 4);
        P.sbufs.tflag = reinterpret_cast<int*>(take(4 * std::max<size_t>(1, P.stiles.size())));
        P.sbufs.cflag = reinterpret_cast<int*>(take(4 * std::max<size_t>(1, P.chunks.size())));
    }
    P.bytes = batched ? off : slot_max;
    P.key = key;
    // pointer patching happens in patch_plan once the workspace exists
    P.chamfer.resize(P.items.size());
    // remember the relative layout inside the items/chunks (stored as offsets)
    for (size_t i = 0; i < P.items.size(); ++i) {
        RItem& it = P.items[i];
        it.warped = reinterpret_cast<void*>(pit[i].warped);
        it.scov = reinterpret_cast<uint8_t*>(pit[i].scov);
        it.dist = reinterpret_cast<float*>(pit[i].dist);
        for (int l = 1; l < kMaxLevels; ++l) {
            it.nd[l] = reinterpret_cast<void*>(pit[i].nd[l]);
            it.wt[l] = reinterpret_cast<void*>(pit[i].wt[l]);
        }
    }
    for (size_t c = 0; c < P.chunks.size(); ++c) {
        RChunk& ch = P.chunks[c];
        ch.boxes = reinterpret_cast<int*>(pch[c].boxes);
        ch.owner = reinterpret_cast<int8_t*>(pch[c].owner);
        for (int l = 1; l < kMaxLevels; ++l) ch.recon[l] = reinterpret_cast<void*>(pch[c].recon[l]);
        ch.lohi = reinterpret_cast<int*>(pch[c].lohi);
        ch.prevl = reinterpret_cast<int*>(pch[c].prevl);
        ch.chosen = reinterpret_cast<int*>(pch[c].chosen);
        ch.back = reinterpret_cast<uint8_t*>(pch[c].back);
        ch.costs = reinterpret_cast<uint8_t*>(pch[c].costs);
        ch.counters = reinterpret_cast<int*>(pch[c].counters);
    }
    // tiles, one stage after another so that each stage's tiles are
    // contiguous across chunks (batched launches) and per chunk (the
    // per-chunk forms); pyramid stages are level-major.
    auto add = [&](int id, int tx0, int tx1, int ty0, int ty1) {
        for (int ty = ty0; ty <= ty1; ++ty)
            for (int tx = tx0; tx <= tx1; ++tx) P.tiles.push_back(RTile{id, tx, ty, 0});
    };
    const int nch = static_cast<int>(P.chunks.size());
    P.warp.assign(nch, {});
    P.own.assign(nch, {});
    for (int c = 0; c < nch; ++c) {
        const RChunk& ch = P.chunks[c];
        P.warp[c].off = static_cast<int>(P.tiles.size());
        for (int k = 0; k < ch.count; ++k) {
            RItem& it = P.items[ch.first + k];
            const int* b = it.lbox[0];
            it.wtile0 = static_cast<int>(P.tiles.size());
            add(ch.first + k, b[0] / kWarpTW, b[2] / kWarpTW, b[1] / kWarpTH, b[3] / kWarpTH);
            it.wtilen = static_cast<int>(P.tiles.size()) - it.wtile0;
        }
        P.warp[c].n = static_cast<int>(P.tiles.size()) - P.warp[c].off;
    }
    for (int c = 0; c < nch; ++c) {
        const RChunk& ch = P.chunks[c];
        P.own[c].off = static_cast<int>(P.tiles.size());
        add(c, 0, (ch.w - 1) / kWarpTW, 0, (ch.h - 1) / kWarpTH);
        P.own[c].n = static_cast<int>(P.tiles.size()) - P.own[c].off;
    }
    P.fin = P.own;
    int maxL = 1;
    for (const auto& ch : P.chunks) maxL = std::max(maxL, ch.levels);
    P.red.assign(nch, std::vector<Plan::Range>(kMaxLevels));
    P.comp.assign(nch, std::vector<Plan::Range>(kMaxLevels));
    for (int l = 0; l + 1 < maxL; ++l)
        for (int c = 0; c < nch; ++c) {
            const RChunk& ch = P.chunks[c];
            P.red[c][l].off = static_cast<int>(P.tiles.size());
            if (l + 1 < ch.levels)
                for (int k = 0; k < ch.count; ++k) {
                    const int* b = P.items[ch.first + k].lbox[l + 1];
                    add(ch.first + k, b[0] / kRedTW, b[2] / kRedTW, b[1] / kRedTH, b[3] / kRedTH);
                }
            P.red[c][l].n = static_cast<int>(P.tiles.size()) - P.red[c][l].off;
        }
    for (int l = 1; l < maxL; ++l)
        for (int c = 0; c < nch; ++c) {
            const RChunk& ch = P.chunks[c];
            P.comp[c][l].off = static_cast<int>(P.tiles.size());
            if (l < ch.levels) add(c, 0, (ch.lw[l] - 1) / kCompTW, 0, (ch.lh[l] - 1) / kCompTH);
            P.comp[c][l].n = static_cast<int>(P.tiles.size()) - P.comp[c][l].off;
        }
    P.sfeat.off = static_cast<int>(P.tiles.size());
    P.sfeat_c.assign(nch, Plan::Range{0, 0});
    for (int c = 0; c < nch; ++c) {
        const RChunk& ch = P.chunks[c];
        P.sfeat_c[c].off = static_cast<int>(P.tiles.size());
        if (ch.sk >= 1) add(c, 0, (ch.lw[ch.sk] - 1) / kCompTW, 0, (ch.lh[ch.sk] - 1) / kCompTH);
        P.sfeat_c[c].n = static_cast<int>(P.tiles.size()) - P.sfeat_c[c].off;
    }
    P.sfeat.n = static_cast<int>(P.tiles.size()) - P.sfeat.off;
    return true;
}

static void patch_plan(Plan& P, uint8_t* base) {
    auto fix = [&](auto*& p) {
        using T = std::remove_reference_t<decltype(p)>;
        p = reinterpret_cast<T>(base + reinterpret_cast<uintptr_t>(p));
    };
    if (P.surr != kSurrNone) {
        fix(P.sbufs.feat); fix(P.sbufs.aux); fix(P.sbufs.out); fix(P.sbufs.tmp0); fix(P.sbufs.tmp1);
        fix(P.sbufs.tflag); fix(P.sbufs.cflag);
    }
    for (auto& it : P.items) {
        fix(it.warped); fix(it.scov); fix(it.dist);
        for (int l = 1; l < kMaxLevels; ++l) { fix(it.nd[l]); fix(it.wt[l]); }
    }
    for (auto& ch : P.chunks) {
        fix(ch.boxes); fix(ch.owner); fix(ch.lohi); fix(ch.prevl); fix(ch.chosen); fix(ch.back); fix(ch.costs); fix(ch.counters);
        for (int l = 1; l < kMaxLevels; ++l) fix(ch.recon[l]);
        for (int k = 0; k < ch.count; ++k) P.items[ch.first + k].box = ch.boxes + 4 * k;
    }
    for (size_t i = 0; i < P.items.size(); ++i) {
        const RItem& it = P.items[i];
        const RChunk& ch = P.chunks[it.chunk];
        // the warp writes the sheared coverage: no row-major copy
        P.chamfer[i] = ChamferWork{nullptr, it.scov, it.dist, it.box, ch.w, ch.h};
    }
}

// Device copies of the tables.
struct DevTables {
    RItem* items = nullptr;
    RChunk* chunks = nullptr;
    RTile* tiles = nullptr;
    ChamferWork* chamfer = nullptr;
    SeamWork<RgbaImg<double, double>>* seam64 = nullptr;
    SeamWork<RgbaImg<float, float>>* seam32 = nullptr;
    SeamWork<MapImg>* seamMap = nullptr;
    SurrImg* simgs = nullptr;
    SurrTile* stiles = nullptr;
    int nst = 0;
    TileMap* tmaps = nullptr;  // per warp tile (precision forms)
    int4* wbox = nullptr;      // per warp tile and wavefront: partial coverage box
    std::vector<std::vector<int>> round_chunks;  // seam round k -> chunk ids
    std::vector<int> round_off;                  // offset of round k in the seam table
};

template <typename Img>
static Img make_img(const RItem& it, const RChunk&) {
    return Img{it.warped};
}
template <>
MapImg make_img<MapImg>(const RItem& it, const RChunk& ch) {
    MapImg m{};
    m.src = it.src; m.fw = it.fw; m.fh = it.fh; m.fmt = it.fmt; m.gain = it.gain;
    // the inverse is filled on the device; build it here with the same formula
    const double* f = it.hfwd;
    const double a = f[0], b = f[1], c = f[2], d = f[3], e = f[4], g5 = f[5], g = f[6], k = f[7], j = f[8];
    const double c0 = e * j - g5 * k, c1 = d * j - g5 * g, c2 = d * k - e * g;
    const double det = (a * c0 - b * c1) + c * c2;
    const double s = 1.0 / det;
    m.h[0] = c0 * s; m.h[1] = (c * k - b * j) * s; m.h[2] = (b * g5 - c * e) * s;
    m.h[3] = (g5 * g - d * j) * s; m.h[4] = (a * j - c * g) * s; m.h[5] = (c * d - a * g5) * s;
    m.h[6] = c2 * s; m.h[7] = (b * g - a * k) * s; m.h[8] = (a * e - b * d) * s;
    m.cx = ch.x; m.cy = ch.y; m.w = ch.w;
    return m;
}

// covShift: the shear of the coverage planes the warp writes for this plan
// (the seam scan reads the sheared coverage).
template <typename Img>
static void seam_table(const Plan& P, std::vector<SeamWork<Img>>& out, DevTables& T, int covShift) {
    size_t rounds = 0;
    for (auto& pr : P.pairs)
        for (const SeamPair& p : pr) rounds = std::max(rounds, static_cast<size_t>(p.round) + 1);
    T.round_chunks.assign(rounds, {});
    T.round_off.assign(rounds, 0);
    // round-major, chunk order within a round (the per-chunk forms launch
    // the contiguous entries of their chunk range)
    for (size_t k = 0; k < rounds; ++k) {
        T.round_off[k] = static_cast<int>(out.size());
        for (size_t c = 0; c < P.chunks.size(); ++c) {
            const RChunk& ch = P.chunks[c];
            const size_t lines = static_cast<size_t>(std::max(ch.w, ch.h));
            for (const SeamPair& p : P.pairs[c]) {
                if (p.round != static_cast<int>(k)) continue;
                const RItem& a = P.items[ch.first + p.i];
                const RItem& b = P.items[ch.first + p.j];
                const size_t sl = static_cast<size_t>(p.slot) * lines;
                SeamWork<Img> s{};
                s.imgI = make_img<Img>(a, ch);
                s.imgJ = make_img<Img>(b, ch);
                s.covI = a.scov; s.covJ = b.scov; s.covShift = covShift;
                s.boxes = ch.boxes; s.ci = p.i; s.cj = p.j; s.w = ch.w;
                s.owner = ch.owner;
                s.lohi = ch.lohi + 2 * sl; s.prevl = ch.prevl + sl; s.chosen = ch.chosen + sl; s.back = ch.back + 64 * sl;
                s.costs = ch.costs + 64 * 8 * sl;
                s.contested = ch.counters + p.i * ch.count + p.j;
                out.push_back(s);
                T.round_chunks[k].push_back(static_cast<int>(c));
            }
        }
    }
}

#define RL(st, NAME, BYTES, ...)                                                                    \
    do {                                                                                            \
        ProfScope ps_(ctx, s, NAME, static_cast<double>(BYTES));                                    \
        __VA_ARGS__;                                                                                \
        hipError_t e_ = hipGetLastError();                                                          \
        if (e_ != hipSuccess) {                                                                     \
            (st).set(I360HIP_KERNEL_LAUNCH_FAILED, "%s launch: %s", NAME, HipName(e_));             \
            return;                                                                                 \
        }                                                                                           \
    } while (0)

// Share of the chip a persistent launch asks for: 8 blocks per CU, divided
// among the chunk groups that run concurrently (set while they are
// enqueued).
static thread_local int g_grid_share = 1;

static unsigned persistent_grid(i360hip_ctx* ctx, int n, int blocks_per_cu = 8) {
    const int cap = std::max(1, ctx->props.multiProcessorCount * blocks_per_cu / g_grid_share);
    return static_cast<unsigned>(std::max(1, std::min(n, cap)));
}

// Which stages of a chunk range enqueue() launches: all of them, the warp
// phase (box init, tile map, warp, box reduce) or the stages after it.
enum EnqueuePart { kPartAll, kPartWarp, kPartRest };

// Launches the render of chunks [c0, c1) on stream s.  Every stage's tiles
// of a chunk range are contiguous, so each stage is one launch over the
// range.
template <int MODE>
static void enqueue(i360hip_ctx* ctx, const Plan& P, const DevTables& T, int c0, int c1, hipStream_t s, Status& st,
                    const SurrRun* sr = nullptr, EnqueuePart part = kPartAll) {
    constexpr int MM = MODE == kModeBatched ? kModePrecision : MODE;
    using M = typename Form<MM>::M;
    using S = typename Form<MM>::S;
    const size_t sS = sizeof(S), sM = sizeof(M);
    auto span = [&](const std::vector<Plan::Range>& v, int& off, int& n) {
        off = v[c0].off;
        n = 0;
        for (int k = c0; k < c1; ++k) n += v[k].n;
    };
    int off, n;
    double area0 = 0;
    for (int k = c0; k < c1; ++k) area0 += static_cast<double>(P.chunks[k].w) * P.chunks[k].h;
    if (part != kPartRest) {
        {
            const int i0 = P.chunks[c0].first;
            const int i1 = P.chunks[c1 - 1].first + P.chunks[c1 - 1].count;
            if (i1 > i0) RL(st, "box_init", 0.0, rbox_init_kernel<<<(i1 - i0 + 63) / 64, 64, 0, s>>>(T.items + i0, T.chunks, i1 - i0));
        }
        // warp
        span(P.warp, off, n);
        if (MM == kModePrecision && n > 0)
            RL(st, "tile_map", 0.0, rtilemap_kernel<<<(n + 255) / 256, 256, 0, s>>>(T.items, T.chunks, T.tiles + off, n, T.tmaps + off));
        RL(st, MODE == kModeFP64 ? "warp_fp64" : (MODE == kModePlain ? "warp_f32" : "warp_f16"),
           static_cast<double>(n) * kWarpTW * kWarpTH * (6.0 + 4 * sS + 1),
           rwarp_kernel<MODE><<<persistent_grid(ctx, n), 256, 0, s>>>(T.items, T.chunks, T.tiles + off, n,
                                                                        T.tmaps + off, T.wbox + 4 * static_cast<size_t>(off),
                                                                        ChamferShear(P.max_w, ctx->tune_strip)));
        {
            const int i0 = P.chunks[c0].first;
            const int i1 = P.chunks[c1 - 1].first + P.chunks[c1 - 1].count;
            if (i1 > i0)
                RL(st, "box_reduce", 0.0, rbox_reduce_kernel<<<std::min(i1 - i0, 1024), 256, 0, s>>>(T.items + i0, i1 - i0, T.wbox));
        }
    }
    if (part == kPartWarp) return;
    // chamfer over the items of the chunk range
    {
        const int i0 = P.chunks[c0].first;
        const int i1 = P.chunks[c1 - 1].first + P.chunks[c1 - 1].count;
        // algorithmic bytes: coverage byte in, distance out (forward), then
        // distance in and out (backward), 13 B per pixel of each item's
        // level-0 support box (the coverage box lies inside it)
        double box_px = 0;
        for (int i = i0; i < i1; ++i) {
            const int* b = P.items[i].lbox[0];
            if (b[2] >= b[0] && b[3] >= b[1]) box_px += static_cast<double>(b[2] - b[0] + 1) * (b[3] - b[1] + 1);
        }
        ProfScope ps_(ctx, s, "chamfer", 13.0 * box_px);
        // the warp wrote the sheared coverage: no shear pre-pass
        if (i1 > i0 && launch_chamfer(T.chamfer + i0, i1 - i0, P.max_w, s, ctx->tune_strip, true) != hipSuccess) {
            st.set(I360HIP_KERNEL_LAUNCH_FAILED, "chamfer launch failed");
            return;
        }
    }
    span(P.own, off, n);
    // algorithmic bytes: each item's distances over its support box (4 B
    // per pixel), the owner byte of every chunk pixel
    double own_box_px = 0;
    for (int i = P.chunks[c0].first; i < P.chunks[c1 - 1].first + P.chunks[c1 - 1].count; ++i) {
        const int* b = P.items[i].lbox[0];
        if (b[2] >= b[0] && b[3] >= b[1]) own_box_px += static_cast<double>(b[2] - b[0] + 1) * (b[3] - b[1] + 1);
    }
    RL(st, "ownership", 4.0 * own_box_px + area0, rowner_kernel<<<persistent_grid(ctx, n, 8), kWarpTW * kWarpTH / kOwnerPix, 0, s>>>(T.items, T.chunks, T.tiles + off, n,
                                                                                   ChamferShear(P.max_w, ctx->tune_strip)));
    // seams: the k-th pair of every chunk in range, in pair order
    {
        for (size_t k = 0; k < T.round_chunks.size(); ++k) {
            const auto& rc = T.round_chunks[k];
            int first = -1, cnt = 0;
            for (size_t q = 0; q < rc.size(); ++q) {
                if (rc[q] < c0 || rc[q] >= c1) continue;
                if (first < 0) first = T.round_off[k] + static_cast<int>(q);
                ++cnt;
            }
            if (cnt == 0) continue;
            int maxLines = 1;
            for (int k2 = c0; k2 < c1; ++k2) maxLines = std::max({maxLines, P.chunks[k2].w, P.chunks[k2].h});
            const dim3 sg((maxLines + kScanLinesPerBlock - 1) / kScanLinesPerBlock, cnt);
            {
                ProfScope ps_(ctx, s, "seam_scan", 0.0);
                if (MM == kModeFP64)
                    seam_scan_table_kernel<RgbaImg<double, double>, double><<<sg, 256, 0, s>>>(T.seam64 + first);
                else if (MM == kModePlain)
                    seam_scan_table_kernel<RgbaImg<float, float>, float><<<sg, 256, 0, s>>>(T.seam32 + first);
                else
                    seam_scan_table_kernel<MapImg, double><<<sg, 256, 0, s>>>(T.seamMap + first);
            }
            {
                ProfScope ps_(ctx, s, "seam_dp", 0.0);
                if (MM == kModeFP64)
                    seam_dp_table_kernel<RgbaImg<double, double>, double><<<cnt, kDpThreads, 0, s>>>(T.seam64 + first);
                else if (MM == kModePlain)
                    seam_dp_table_kernel<RgbaImg<float, float>, float><<<cnt, kDpThreads, 0, s>>>(T.seam32 + first);
                else
                    seam_dp_table_kernel<MapImg, double><<<cnt, kDpThreads, 0, s>>>(T.seamMap + first);
            }
            {
                ProfScope ps_(ctx, s, "seam_reassign", 0.0);
                if (MM == kModeFP64)
                    seam_reassign_table_kernel<RgbaImg<double, double>><<<sg, 256, 0, s>>>(T.seam64 + first);
                else if (MM == kModePlain)
                    seam_reassign_table_kernel<RgbaImg<float, float>><<<sg, 256, 0, s>>>(T.seam32 + first);
                else
                    seam_reassign_table_kernel<MapImg><<<sg, 256, 0, s>>>(T.seamMap + first);
            }
            if (hipGetLastError() != hipSuccess) {
                st.set(I360HIP_KERNEL_LAUNCH_FAILED, "seam launch failed");
                return;
            }
        }
    }
    // pyramids (with the surrogate: the levels up to K first, the network,
    // the levels above K only for chunks that fall back, then the rest)
    int maxL = 1;
    for (int k = c0; k < c1; ++k) maxL = std::max(maxL, P.chunks[k].levels);
    const bool surr = sr && sr->mode == kSurrRender;
    const bool exporting = sr && sr->mode == kSurrExport;
    const int* cflag = sr ? sr->bufs.cflag : nullptr;
    auto reduces = [&](int phase) {
        for (int l = 0; l + 1 < maxL; ++l) {
            off = P.red[c0][l].off;
            n = 0;
            for (int k = c0; k < c1; ++k) n += P.red[k][l].n;
            if (n == 0) continue;
            // per level-(l+1) tile pixel: four level-l pixels (RGBA, and the
            // weight: at level 0 the owner byte) in, RGBA and weight out
            const double bytes = static_cast<double>(n) * kRedTW * kRedTH *
                                 (4.0 * (4 * sS + (l == 0 ? 1.0 : sS)) + (4 * sS + sS));
            if (l == 0 && MM == kModePrecision)
                RL(st, "reduce_l0", bytes, rreduce_s_kernel<MODE, true><<<persistent_grid(ctx, n), 256, 0, s>>>(T.items, T.chunks, T.tiles + off, n, 0, phase, cflag));
            else if (MM == kModePrecision)
                RL(st, "reduce", bytes, rreduce_s_kernel<MODE, false><<<persistent_grid(ctx, n), 256, 0, s>>>(T.items, T.chunks, T.tiles + off, n, l, phase, cflag));
            else if (l == 0)
                RL(st, "reduce_l0", bytes, rreduce_kernel<MODE, true><<<persistent_grid(ctx, n), 256, 0, s>>>(T.items, T.chunks, T.tiles + off, n, l, phase, cflag));
            else
                RL(st, "reduce", bytes, rreduce_kernel<MODE, false><<<persistent_grid(ctx, n), 256, 0, s>>>(T.items, T.chunks, T.tiles + off, n, l, phase, cflag));
        }
    };
    auto composes = [&](int phase) {
        for (int l = maxL - 1; l >= 1; --l) {
            off = P.comp[c0][l].off;
            n = 0;
            for (int k = c0; k < c1; ++k) n += P.comp[k][l].n;
            if (n == 0) continue;
            const double bytes = static_cast<double>(n) * kCompTW * kCompTH * (4 * sM + 2.0 * (4 * sS + sS));
            RL(st, "compose_level", bytes,
               rcompose_kernel<MODE><<<persistent_grid(ctx, n), 256, 0, s>>>(T.items, T.chunks, T.tiles + off, n, l, phase, cflag));
        }
    };
    // the surrogate stages of chunks [c0, c1): their feature tiles and
    // network tiles are contiguous sub-ranges (chunk order)
    int fo = 0, fn = 0, so = 0, sn = 0, spx = 0;
    if (sr && !P.sfeat_c.empty() && !P.stile_c.empty()) {
        span(P.sfeat_c, fo, fn);
        span(P.stile_c, so, sn);
        for (int k = c0; k < c1; ++k)
            if (P.chunks[k].sk >= 1) spx += P.chunks[k].lw[P.chunks[k].sk] * P.chunks[k].lh[P.chunks[k].sk];
    }
    if (surr) {
        const int nc = c1 - c0;
        RL(st, "surrogate_flags", 0.0,
           rscflag_init_kernel<<<(nc + 255) / 256, 256, 0, s>>>(T.chunks + c0, nc, sr->bufs.cflag + c0));
    }
    reduces(surr ? 1 : 0);
    if constexpr (MODE == kModeBatched) if (sr && fn > 0) {
        const double px = static_cast<double>(spx);
        RL(st, "surrogate_features", px * 64.0,
           rsfeat_kernel<MODE><<<persistent_grid(ctx, fn), 256, 0, s>>>(T.items, T.chunks, T.tiles + fo, fn, sr->bufs));
        RL(st, "surrogate_blur", px * 80.0,
           rsblur_kernel<<<persistent_grid(ctx, fn), 256, 0, s>>>(T.chunks, T.tiles + fo, fn, sr->bufs, 0));
        RL(st, "surrogate_blur", px * 96.0,
           rsblur_kernel<<<persistent_grid(ctx, fn), 256, 0, s>>>(T.chunks, T.tiles + fo, fn, sr->bufs, 1));
        if (surr && sn > 0) {
            {
                ProfScope ps_(ctx, s, "surrogate_network", 0.0);
                if (surr_forward_tables(T.simgs, T.stiles + so, sn, sr->weights, sr->biases, sr->bufs.feat, sr->bufs.out,
                                        sr->bufs.tmp0, sr->bufs.tmp1, s) != hipSuccess) {
                    st.set(I360HIP_KERNEL_LAUNCH_FAILED, "surrogate network launch failed");
                    return;
                }
            }
            RL(st, "surrogate_flags", px * 16.0,
               rsflag_kernel<<<sn, 256, 0, s>>>(T.simgs, T.stiles, so, sr->bufs, sr->threshold, sr->max_threshold));
        }
    }
    if (surr) {
        reduces(2);
        composes(2);
        if constexpr (MODE == kModeBatched) if (fn > 0)
            RL(st, "surrogate_apply", static_cast<double>(spx) * 48.0,
               rsapply_kernel<MODE><<<persistent_grid(ctx, fn), 256, 0, s>>>(T.chunks, T.tiles + fo, fn, sr->bufs));
        composes(1);
    } else {
        composes(0);
    }
    if constexpr (MODE == kModeBatched) if (exporting && fn > 0)
        RL(st, "surrogate_export", static_cast<double>(spx) * 28.0,
           rsexport_kernel<MODE><<<persistent_grid(ctx, fn), 256, 0, s>>>(T.chunks, T.tiles + fo, fn, sr->bufs));
    if (exporting) return;  // the export writes no chunk output
    span(P.fin, off, n);
    const double out_bytes = (c1 > c0 && P.chunks[c0].out16) ? 6.0 : 12.0;  // RGB48 or FP32 RGB per pixel
    // per chunk pixel: the owner byte, the owner's RGBA, a quarter texel of
    // the level-1 reconstruction (RGBA, maths type) and of the owner's
    // level-1 numerator / denominator, the output
    RL(st, "final_level0", area0 * (1.0 + 4 * sS + (4 * sM + 4 * sS) / 4.0 + out_bytes),
       rfinal_kernel<MODE><<<persistent_grid(ctx, n, 8), 256, 0, s>>>(T.items, T.chunks, T.tiles + off, n));
}

struct RenderCache {
    uint64_t key = 0;
    uint64_t call = 0;  // call_key of the call that captured the graph
    int mode = -1;
    Plan plan;
    DevTables tables;
    hipGraphExec_t graph = nullptr;
    hipStream_t graph_stream = nullptr;
    // chunk groups of the batched form: auxiliary streams and their events
    hipEvent_t fork = nullptr;
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> forks;  // per auxiliary stream (staggered groups)
    std::vector<hipEvent_t> join;
    // recorded after the last enqueued render (any stream): the workspace
    // and tables are not rewritten before it has completed
    hipEvent_t done = nullptr;
};

static RenderCache& cache_of(i360hip_ctx* ctx) {
    if (!ctx->render_cache) ctx->render_cache = new RenderCache();
    return *static_cast<RenderCache*>(ctx->render_cache);
}

void render_release(i360hip_ctx* ctx) {
    if (!ctx->render_cache) return;
    auto* rc = static_cast<RenderCache*>(ctx->render_cache);
    if (rc->graph) (void)hipGraphExecDestroy(rc->graph);
    for (auto a : rc->aux) {
        (void)hipStreamSynchronize(a);
        (void)hipStreamDestroy(a);
    }
    for (auto e : rc->join) (void)hipEventDestroy(e);
    for (auto e : rc->forks) (void)hipEventDestroy(e);
    if (rc->fork) (void)hipEventDestroy(rc->fork);
    if (rc->done) {
        (void)hipEventSynchronize(rc->done);
        (void)hipEventDestroy(rc->done);
    }
    delete rc;
    ctx->render_cache = nullptr;
}

// A render enqueued on one stream may still be reading the workspace, the
// tables or the surrogate parameters when a later call on another stream
// rewrites them (a new plan, new parameters): every render records `done` on
// its stream, and a call that rewrites waits for it first.
static void mark_done(RenderCache& rc, hipStream_t s, Status& st) {
    if (!rc.done && hipEventCreateWithFlags(&rc.done, hipEventDisableTiming) != hipSuccess) {
        rc.done = nullptr;
        st.set(I360HIP_DEVICE_ERROR, "render completion event creation failed");
        return;
    }
    if (hipEventRecord(rc.done, s) != hipSuccess) st.set(I360HIP_DEVICE_ERROR, "render completion record failed");
}
static bool wait_renders(const RenderCache& rc) { return !rc.done || hipEventSynchronize(rc.done) == hipSuccess; }

// Uploads the tables of `P` into the table slot; fills T.
static void upload_tables(i360hip_ctx* ctx, Plan& P, DevTables& T, hipStream_t s, Status& st, int slot) {
    std::vector<SeamWork<RgbaImg<double, double>>> s64;
    std::vector<SeamWork<RgbaImg<float, float>>> s32;
    std::vector<SeamWork<MapImg>> s16;
    T = DevTables{};
    const int covShift = ChamferShear(P.max_w, ctx->tune_strip);  // as the warp's launch
    if (P.mode == kModeFP64) seam_table(P, s64, T, covShift);
    else if (P.mode == kModePlain) seam_table(P, s32, T, covShift);
    else seam_table(P, s16, T, covShift);
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += A256(b); return o; };
    const size_t oI = take(sizeof(RItem) * std::max<size_t>(1, P.items.size()));
    const size_t oC = take(sizeof(RChunk) * std::max<size_t>(1, P.chunks.size()));
    const size_t oT = take(sizeof(RTile) * std::max<size_t>(1, P.tiles.size()));
    const size_t oW = take(sizeof(ChamferWork) * std::max<size_t>(1, P.chamfer.size()));
    const size_t oS = take(std::max({sizeof(s64[0]) * s64.size(), sizeof(s32[0]) * s32.size(),
                                     sizeof(s16[0]) * s16.size(), size_t(256)}));
    const size_t oSI = take(sizeof(SurrImg) * std::max<size_t>(1, P.simgs.size()));
    const size_t oST = take(sizeof(SurrTile) * std::max<size_t>(1, P.stiles.size()));
    size_t nwarp = 0;
    for (const auto& r : P.warp) nwarp += r.n;
    const size_t upload = off;  // device-only arrays follow the uploaded tables
    const size_t oTM = take(sizeof(TileMap) * std::max<size_t>(1, nwarp));
    const size_t oWB = take(sizeof(int4) * 4 * std::max<size_t>(1, nwarp));
    std::vector<uint8_t> host(upload);
    std::memcpy(host.data() + oI, P.items.data(), sizeof(RItem) * P.items.size());
    std::memcpy(host.data() + oC, P.chunks.data(), sizeof(RChunk) * P.chunks.size());
    std::memcpy(host.data() + oT, P.tiles.data(), sizeof(RTile) * P.tiles.size());
    std::memcpy(host.data() + oW, P.chamfer.data(), sizeof(ChamferWork) * P.chamfer.size());
    if (!s64.empty()) std::memcpy(host.data() + oS, s64.data(), sizeof(s64[0]) * s64.size());
    if (!s32.empty()) std::memcpy(host.data() + oS, s32.data(), sizeof(s32[0]) * s32.size());
    if (!s16.empty()) std::memcpy(host.data() + oS, s16.data(), sizeof(s16[0]) * s16.size());
    if (!P.simgs.empty()) std::memcpy(host.data() + oSI, P.simgs.data(), sizeof(SurrImg) * P.simgs.size());
    if (!P.stiles.empty()) std::memcpy(host.data() + oST, P.stiles.data(), sizeof(SurrTile) * P.stiles.size());
    auto* dev = static_cast<uint8_t*>(ctx_scratch(ctx, slot, off, st));
    if (!st.ok()) return;
    if (hipMemcpyAsync(dev, host.data(), upload, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        st.set(I360HIP_DEVICE_ERROR, "render table upload failed");
        return;
    }
    T.items = reinterpret_cast<RItem*>(dev + oI);
    T.chunks = reinterpret_cast<RChunk*>(dev + oC);
    T.tiles = reinterpret_cast<RTile*>(dev + oT);
    T.chamfer = reinterpret_cast<ChamferWork*>(dev + oW);
    T.seam64 = reinterpret_cast<decltype(T.seam64)>(dev + oS);
    T.seam32 = reinterpret_cast<decltype(T.seam32)>(dev + oS);
    T.seamMap = reinterpret_cast<decltype(T.seamMap)>(dev + oS);
    T.simgs = reinterpret_cast<SurrImg*>(dev + oSI);
    T.stiles = reinterpret_cast<SurrTile*>(dev + oST);
    T.nst = static_cast<int>(P.stiles.size());
    T.tmaps = reinterpret_cast<TileMap*>(dev + oTM);
    T.wbox = reinterpret_cast<int4*>(dev + oWB);
}

// Chunk groups of the batched form: contiguous chunk ranges of about equal
// output area, each rendered on its own stream (the context's auxiliary
// streams, forked from and joined back into `s`, also under graph capture),
// so the latency-bound stages of one group (chamfer, seam DP) overlap the
// bandwidth-bound stages of the others.  The count comes from the
// autotuner (i360hip_render_autotune), default 4; I360HIP_RENDER_GROUPS
// overrides it (1 = one launch sequence over every chunk).
constexpr int kMaxGroups = 8;

static int render_groups(const i360hip_ctx* ctx, int nch) {
    int g = ctx->tune_groups > 0 ? ctx->tune_groups : 4;
    if (const char* e = std::getenv("I360HIP_RENDER_GROUPS")) g = std::atoi(e);
    return std::max(1, std::min({g, kMaxGroups, nch}));
}

static bool aux_streams(i360hip_ctx* ctx, int n) {
    RenderCache& rc = cache_of(ctx);
    if (!rc.fork) {
        if (hipEventCreateWithFlags(&rc.fork, hipEventDisableTiming) != hipSuccess) return false;
    }
    while (static_cast<int>(rc.aux.size()) < n) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr, fk = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return false;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipStreamDestroy(st);
            return false;
        }
        if (hipEventCreateWithFlags(&fk, hipEventDisableTiming) != hipSuccess) {
            (void)hipEventDestroy(ev);
            (void)hipStreamDestroy(st);
            return false;
        }
        rc.aux.push_back(st);
        rc.join.push_back(ev);
        rc.forks.push_back(fk);
    }
    return true;
}

// Staggered chunk groups: the groups' warp phases run one after another on
// the origin stream, and each group's remaining stages fork off onto its
// own stream as soon as its warp phase is done, so the groups reach their
// latency-bound stages (chamfer, seams) at different times and each one's
// serial chain overlaps the others' bandwidth-bound stages.  Off: every
// group forks at the start and all groups run their stages side by side.
// I360HIP_RENDER_STAGGER=0/1 overrides the default (on).
static bool render_stagger() {
    if (const char* e = std::getenv("I360HIP_RENDER_STAGGER")) return std::atoi(e) != 0;
    return true;
}

static void enqueue_groups(i360hip_ctx* ctx, const Plan& P, const DevTables& T, hipStream_t s, Status& st,
                           const SurrRun* sr = nullptr) {
    const int nch = static_cast<int>(P.chunks.size());
    const int G = render_groups(ctx, nch);
    if (G <= 1 || !aux_streams(ctx, G - 1)) {
        enqueue<kModeBatched>(ctx, P, T, 0, nch, s, st, sr);
        return;
    }
    RenderCache& rc = cache_of(ctx);
    double total = 0;
    for (const auto& c : P.chunks) total += static_cast<double>(c.w) * c.h;
    // Split points by output area.  Staggered, the step ends with the last
    // group's pyramids (after every warp phase and its own serial chain),
    // and no group's pyramids start before the first group's chain ends
    // (after the first warp phase): so the first and the last group get two
    // shares of the area and every other group three (20/30/30/20 % for
    // four groups: gigapixel step 30.45-31.27 -> 29.35-29.66 ms against
    // equal shares, three runs each on one box).  Side by side, equal
    // shares.
    const bool stagger = render_stagger();
    std::vector<double> share_of(G, 1.0);
    if (stagger && G > 2)
        for (int g = 0; g < G; ++g) share_of[g] = (g == 0 || g == G - 1) ? 2.0 : 3.0;
    double wsum = 0;
    for (double v : share_of) wsum += v;
    std::vector<int> cut{0};
    double acc = 0, bound = total * share_of[0] / wsum;
    for (int c = 0; c < nch && static_cast<int>(cut.size()) < G; ++c) {
        acc += static_cast<double>(P.chunks[c].w) * P.chunks[c].h;
        if (acc >= bound && c + 1 < nch) {
            cut.push_back(c + 1);
            if (static_cast<int>(cut.size()) < G) bound += total * share_of[cut.size() - 1] / wsum;
        }
    }
    cut.push_back(nch);
    const int ng = static_cast<int>(cut.size()) - 1;
    int share = ctx->tune_share > 0 ? ctx->tune_share : 1;
    if (const char* e = std::getenv("I360HIP_GRID_SHARE")) share = std::max(1, std::atoi(e));
    if (stagger) {
        // Each fork event is waited on right after it is recorded, before
        // anything more is enqueued on the origin stream.
        g_grid_share = share;
        int forked = 0;
        for (int g = 0; g < ng && st.ok(); ++g) {
            enqueue<kModeBatched>(ctx, P, T, cut[g], cut[g + 1], s, st, sr, kPartWarp);
            if (!st.ok()) break;
            if (g + 1 == ng) {
                enqueue<kModeBatched>(ctx, P, T, cut[g], cut[g + 1], s, st, sr, kPartRest);
                break;
            }
            if (hipEventRecord(rc.forks[g], s) != hipSuccess || hipStreamWaitEvent(rc.aux[g], rc.forks[g], 0) != hipSuccess) {
                st.set(I360HIP_DEVICE_ERROR, "render fork failed");
                break;
            }
            ++forked;
            enqueue<kModeBatched>(ctx, P, T, cut[g], cut[g + 1], rc.aux[g], st, sr, kPartRest);
        }
        g_grid_share = 1;
        for (int g = 0; g < forked; ++g)
            if (hipEventRecord(rc.join[g], rc.aux[g]) != hipSuccess || hipStreamWaitEvent(s, rc.join[g], 0) != hipSuccess) {
                if (st.ok()) st.set(I360HIP_DEVICE_ERROR, "render join failed");
            }
        return;
    }
    if (hipEventRecord(rc.fork, s) != hipSuccess) {
        st.set(I360HIP_DEVICE_ERROR, "render fork failed");
        return;
    }
    for (int g = 1; g < ng; ++g)
        if (hipStreamWaitEvent(rc.aux[g - 1], rc.fork, 0) != hipSuccess) {
            st.set(I360HIP_DEVICE_ERROR, "render fork failed");
            return;
        }
    g_grid_share = share;
    for (int g = 0; g < ng && st.ok(); ++g)
        enqueue<kModeBatched>(ctx, P, T, cut[g], cut[g + 1], g == 0 ? s : rc.aux[g - 1], st, sr);
    g_grid_share = 1;
    for (int g = 1; g < ng; ++g)
        if (hipEventRecord(rc.join[g - 1], rc.aux[g - 1]) != hipSuccess ||
            hipStreamWaitEvent(s, rc.join[g - 1], 0) != hipSuccess) {
            if (st.ok()) st.set(I360HIP_DEVICE_ERROR, "render join failed");
        }
}

// Hash of everything a render call's plan depends on, read straight from
// the call's arguments: frames (pixels, size, format, gain, homography),
// chunks (rectangle, contributor range, outputs), the contributor lists
// (as far as the planner reads them), bands, form and launch tuning.  A
// repeated call with an equal hash replays its captured graph without
// planning again (planning a gigapixel canvas costs ~20 ms of host time).
static uint64_t call_key(const i360hip_ctx* ctx, const i360hip_frame* frames, int nf, const i360hip_chunk* chunks,
                         int nc, const int* contrib, int bands, int api_mode) {
    uint64_t k = mix(0x1a2b3c4d5e6f7788ull, static_cast<uint64_t>(api_mode));
    k = mix(k, static_cast<uint64_t>(bands));
    k = mix(k, static_cast<uint64_t>(nf) * 1000003ull + static_cast<uint64_t>(nc));
    k = mix(k, static_cast<uint64_t>(ctx->tune_groups * 1000003 + ctx->tune_strip * 101 + ctx->tune_share));
    auto bits = [](auto v) {
        uint64_t b = 0;
        std::memcpy(&b, &v, sizeof(v));
        return b;
    };
    for (int f = 0; f < nf; ++f) {
        const i360hip_frame& fr = frames[f];
        k = mix(k, reinterpret_cast<uintptr_t>(fr.pixels));
        k = mix(k, static_cast<uint64_t>(fr.width) * 65537ull + static_cast<uint64_t>(fr.height) * 3ull + fr.format);
        k = mix(k, bits(fr.gain));
        for (int q = 0; q < 9; ++q) k = mix(k, bits(fr.h[q]));
    }
    for (int c = 0; c < nc; ++c) {
        const i360hip_chunk& cc = chunks[c];
        k = mix(k, static_cast<uint64_t>(cc.x) * 1000003ull + static_cast<uint64_t>(cc.y));
        k = mix(k, static_cast<uint64_t>(cc.width) * 1000003ull + static_cast<uint64_t>(cc.height));
        k = mix(k, static_cast<uint64_t>(cc.first) * 1000003ull + static_cast<uint64_t>(cc.count));
        k = mix(k, reinterpret_cast<uintptr_t>(cc.out));
        k = mix(k, reinterpret_cast<uintptr_t>(cc.owner));
        if (cc.first < 0 || cc.count < 0) continue;  // the planner rejects it before reading
        const int cap = std::min(cc.count, I360HIP_MAX_CONTRIBUTORS);
        for (int q = 0; q < cap; ++q) k = mix(k, static_cast<uint64_t>(static_cast<uint32_t>(contrib[cc.first + q])));
    }
    return k;
}

static void run_render(i360hip_ctx* ctx, const i360hip_frame* frames, int nf, const i360hip_chunk* chunks, int nc,
                       const int* contrib, int bands, int api_mode, hipStream_t s, Status& st,
                       float* export_feat = nullptr, float* export_target = nullptr) {
    RenderCache& rc = cache_of(ctx);
    const bool out16 = (api_mode & I360HIP_RENDER_OUT_RGB48) != 0;
    const int form = api_mode & ~I360HIP_RENDER_OUT_RGB48;
    const int surr = export_feat ? kSurrExport : (form == I360HIP_RENDER_SURROGATE ? kSurrRender : kSurrNone);
    const int mode = surr != kSurrNone ? kModeBatched : form;
    if (surr == kSurrRender && !ctx->surr_weights) {
        st.set(I360HIP_INVALID_PARAM, "the SURROGATE form needs parameters (i360hip_surrogate_set)");
        return;
    }
    const bool graphable = mode == kModeBatched && surr != kSurrExport;
    const uint64_t ck = graphable ? call_key(ctx, frames, nf, chunks, nc, contrib, bands, api_mode) : 0;
    if (graphable && !ctx->profiling && rc.graph && rc.mode == api_mode && rc.call == ck) {
        if (hipGraphLaunch(rc.graph, s) != hipSuccess) st.set(I360HIP_KERNEL_LAUNCH_FAILED, "graph launch failed");
        else mark_done(rc, s, st);
        return;
    }
    Plan P;
    // The batched form plans the chunks lightest first (fewest contributors,
    // then smallest area; canvas order on ties).  Each chunk group's serial
    // chain (chamfer, seam rounds) is as long as its heaviest chunk's, and
    // no group's pyramids start before the first group's chain ends: with
    // the heavy chunks gathered in the last group, the earlier groups' chains
    // end sooner (gigapixel step 29.6-30.3 -> 28.8-29.1 ms, canvas order
    // against this order, three runs each).  Outputs go to each chunk's own
    // buffers, so the order changes no result.
    std::vector<i360hip_chunk> order;
    std::vector<int> idx;
    const i360hip_chunk* plan_chunks = chunks;
    if (graphable && surr != kSurrExport && nc > 2) {
        idx.resize(nc);
        for (int c = 0; c < nc; ++c) idx[c] = c;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
            if (chunks[a].count != chunks[b].count) return chunks[a].count < chunks[b].count;
            return static_cast<int64_t>(chunks[a].width) * chunks[a].height <
                   static_cast<int64_t>(chunks[b].width) * chunks[b].height;
        });
        order.reserve(nc);
        for (int c : idx) order.push_back(chunks[c]);
        plan_chunks = order.data();
    }
    if (!build_plan(frames, nf, plan_chunks, nc, contrib, bands, mode, P, st, surr, out16, idx.empty() ? nullptr : idx.data()))
        return;
    SurrRun sr;
    sr.mode = surr;
    sr.weights = ctx->surr_weights;
    sr.biases = ctx->surr_biases;
    sr.threshold = ctx->surr_threshold;
    sr.max_threshold = ctx->surr_max_threshold;
    P.key = mix(P.key, static_cast<uint64_t>(ctx->tune_groups * 1000003 + ctx->tune_strip * 101 + ctx->tune_share));
    const bool reuse = !ctx->profiling && rc.graph && rc.key == P.key && rc.mode == api_mode && graphable;
    if (reuse) {  // other arguments, the same plan
        rc.call = ck;
        if (hipGraphLaunch(rc.graph, s) != hipSuccess) st.set(I360HIP_KERNEL_LAUNCH_FAILED, "graph launch failed");
        else mark_done(rc, s, st);
        return;
    }
    if (!wait_renders(rc)) {
        st.set(I360HIP_DEVICE_ERROR, "waiting for the previous render failed");
        return;
    }
    if (rc.graph) {
        (void)hipGraphExecDestroy(rc.graph);
        rc.graph = nullptr;
    }
    auto* base = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotRender, std::max<size_t>(P.bytes, 256), st));
    if (!st.ok()) return;
    patch_plan(P, base);
    upload_tables(ctx, P, rc.tables, s, st, kSlotRenderTables);
    if (!st.ok()) return;
    rc.plan = std::move(P);
    rc.key = rc.plan.key;
    rc.mode = api_mode;
    const Plan& Q = rc.plan;
    const DevTables& T = rc.tables;
    sr.bufs = Q.sbufs;
    sr.bufs.target = export_target;
    const int ni = static_cast<int>(Q.items.size());
    auto body = [&](hipStream_t ss) {
        if (ni > 0) rmap_kernel<<<(ni + 63) / 64, 64, 0, ss>>>(T.items, ni);
        switch (mode) {
            case kModeFP64:
                for (int c = 0; c < static_cast<int>(Q.chunks.size()) && st.ok(); ++c) enqueue<kModeFP64>(ctx, Q, T, c, c + 1, ss, st);
                break;
            case kModePlain:
                for (int c = 0; c < static_cast<int>(Q.chunks.size()) && st.ok(); ++c) enqueue<kModePlain>(ctx, Q, T, c, c + 1, ss, st);
                break;
            case kModePrecision:
                for (int c = 0; c < static_cast<int>(Q.chunks.size()) && st.ok(); ++c) enqueue<kModePrecision>(ctx, Q, T, c, c + 1, ss, st);
                break;
            default:
                if (Q.chunks.empty()) break;
                if (surr == kSurrExport) {
                    enqueue<kModeBatched>(ctx, Q, T, 0, static_cast<int>(Q.chunks.size()), ss, st, &sr);
                    break;
                }
                if (surr == kSurrRender) {  // chunk groups on concurrent streams, as the batched form
                    enqueue_groups(ctx, Q, T, ss, st, &sr);
                    break;
                }
                enqueue_groups(ctx, Q, T, ss, st);
                break;
        }
    };
    if (surr == kSurrExport) {
        body(s);
        if (st.ok() && Q.snpix > 0 &&
            hipMemcpyAsync(export_feat, Q.sbufs.feat, sizeof(float) * kSurrIn * static_cast<size_t>(Q.snpix),
                           hipMemcpyDeviceToDevice, s) != hipSuccess)
            st.set(I360HIP_DEVICE_ERROR, "surrogate feature copy failed");
        mark_done(rc, s, st);
        return;
    }
    if (!graphable || ctx->profiling) {
        body(s);
        mark_done(rc, s, st);
        return;
    }
    // batched: capture once, replay on later identical calls
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        body(s);
        mark_done(rc, s, st);
        return;
    }
    body(s);
    if (hipStreamEndCapture(s, &g) != hipSuccess || !st.ok()) {
        if (g) (void)hipGraphDestroy(g);
        if (st.ok()) st.set(I360HIP_DEVICE_ERROR, "graph capture failed");
        return;
    }
    hipGraphExec_t ge = nullptr;
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        st.set(I360HIP_DEVICE_ERROR, "graph instantiate failed");
        return;
    }
    (void)hipGraphDestroy(g);
    // Upload the executable graph's device-side state now: replays of an
    // uploaded graph reach the queues sooner (measured: four-group
    // gigapixel step 36.8 -> 36.3 ms).
    if (hipError_t e = hipGraphUpload(ge, s); e != hipSuccess) {
        (void)hipGraphExecDestroy(ge);
        st.set(CodeOf(e), "hipGraphUpload failed: %s", HipName(e));
        return;
    }
    rc.graph = ge;
    rc.call = ck;
    if (hipGraphLaunch(ge, s) != hipSuccess) st.set(I360HIP_KERNEL_LAUNCH_FAILED, "graph launch failed");
    else mark_done(rc, s, st);
}

}  // namespace i360

using namespace i360;

extern "C" {

int i360hip_render_chunks(i360hip_ctx* ctx, const i360hip_frame* frames, int num_frames,
                          const i360hip_chunk* chunks, int num_chunks, const int* contributors, int num_bands,
                          int mode, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if ((!frames && num_frames > 0) || (!chunks && num_chunks > 0) || (!contributors && num_chunks > 0)) {
        set_error(ctx, "Null pointer argument.");
        return I360HIP_INVALID_PARAM;
    }
    const int form = mode & ~I360HIP_RENDER_OUT_RGB48;
    if (num_bands < 1 || num_bands > I360HIP_MAX_BANDS || form < I360HIP_RENDER_FP64 || form > I360HIP_RENDER_SURROGATE) {
        set_error(ctx, "numBands must be 1-12 and mode a known form.");
        return I360HIP_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return I360HIP_DEVICE_ERROR;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    run_render(ctx, frames, num_frames, chunks, num_chunks, contributors, num_bands, mode, s, st);
    if (!st.ok()) set_error(ctx, st.msg);
    return st.code;
}

int i360hip_render_map_error(i360hip_ctx* ctx, const i360hip_frame* frames, int num_frames,
                             const i360hip_chunk* chunks, int num_chunks, const int* contributors,
                             double* max_error_px) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!max_error_px || (!frames && num_frames > 0) || (!chunks && num_chunks > 0)) return I360HIP_INVALID_PARAM;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return I360HIP_DEVICE_ERROR;
    Plan P;
    if (!build_plan(frames, num_frames, chunks, num_chunks, contributors, 1, kModePrecision, P, st)) {
        set_error(ctx, st.msg);
        return st.code;
    }
    // only the item table and warp tiles are used; no workspace needed
    for (auto& it : P.items) { it.warped = nullptr; it.scov = nullptr; it.dist = nullptr; }
    DevTables T;
    upload_tables(ctx, P, T, ctx->stream, st, kSlotRender + 3);
    auto* bits = static_cast<unsigned int*>(ctx_scratch(ctx, kSlotRender + 2, 256, st));
    if (!st.ok()) {
        set_error(ctx, st.msg);
        return st.code;
    }
    (void)hipMemsetAsync(bits, 0, sizeof(unsigned int), ctx->stream);
    const int ni = static_cast<int>(P.items.size());
    if (ni > 0) rmap_kernel<<<(ni + 63) / 64, 64, 0, ctx->stream>>>(T.items, ni);
    const int n = static_cast<int>(P.tiles.size());
    // warp tiles come first per chunk; all warp ranges together
    int total = 0;
    for (auto& r : P.warp) total += r.n;
    (void)total;
    for (auto& r : P.warp)
        if (r.n > 0)
            rmap_error_kernel<<<persistent_grid(ctx, r.n), 256, 0, ctx->stream>>>(T.items, T.chunks, T.tiles + r.off, r.n, bits);
    (void)n;
    unsigned int hb = 0;
    if (hipMemcpyAsync(&hb, bits, sizeof(hb), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return I360HIP_DEVICE_ERROR;
    float f;
    std::memcpy(&f, &hb, sizeof(f));
    *max_error_px = static_cast<double>(f);
    return I360HIP_SUCCESS;
}

}  // extern "C"

extern "C" {

int i360hip_surrogate_set(i360hip_ctx* ctx, const uint16_t* weights, const float* biases, float threshold_mse) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!weights || !biases || !(threshold_mse >= 0.0f)) {
        set_error(ctx, "Null pointer argument or negative threshold.");
        return I360HIP_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return I360HIP_DEVICE_ERROR;
    constexpr int kSlotSurrParams = 18;
    if (!wait_renders(cache_of(ctx))) {  // an enqueued SURROGATE render may still read the parameters
        set_error(ctx, "waiting for the previous render failed");
        return I360HIP_DEVICE_ERROR;
    }
    const size_t wb = A256(2 * static_cast<size_t>(I360HIP_SURR_WEIGHTS));
    auto* dev = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotSurrParams, wb + 4 * I360HIP_SURR_BIASES, st));
    if (!st.ok()) {
        set_error(ctx, st.msg);
        return st.code;
    }
    if (hipMemcpyAsync(dev, weights, 2 * static_cast<size_t>(I360HIP_SURR_WEIGHTS), hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dev + wb, biases, 4 * static_cast<size_t>(I360HIP_SURR_BIASES), hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        set_error(ctx, "surrogate parameter copy failed");
        return I360HIP_DEVICE_ERROR;
    }
    ctx->surr_weights = reinterpret_cast<const uint16_t*>(dev);
    ctx->surr_biases = reinterpret_cast<const float*>(dev + wb);
    if (ctx->surr_threshold != threshold_mse) {
        // the threshold is a kernel argument of a captured graph
        RenderCache& rc = cache_of(ctx);
        if (rc.graph) (void)hipGraphExecDestroy(rc.graph);
        rc.graph = nullptr;
        ctx->surr_threshold = threshold_mse;
    }
    return I360HIP_SUCCESS;
}

int i360hip_surrogate_set_max_error(i360hip_ctx* ctx, float max_sq_err) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!(max_sq_err >= 0.0f) || !std::isfinite(max_sq_err)) {
        set_error(ctx, "Surrogate max error bound must be finite and non-negative.");
        return I360HIP_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->surr_max_threshold != max_sq_err) {
        // the bound is a kernel argument of a captured graph
        if (hipSetDevice(ctx->device) != hipSuccess) return I360HIP_DEVICE_ERROR;
        RenderCache& rc = cache_of(ctx);
        if (!wait_renders(rc)) {
            set_error(ctx, "waiting for the previous render failed");
            return I360HIP_DEVICE_ERROR;
        }
        if (rc.graph) (void)hipGraphExecDestroy(rc.graph);
        rc.graph = nullptr;
        ctx->surr_max_threshold = max_sq_err;
    }
    return I360HIP_SUCCESS;
}

int i360hip_surrogate_chunk_level(int width, int height, int num_bands, int* level, int* level_width,
                                  int* level_height) {
    if (!level || !level_width || !level_height || width <= 0 || height <= 0 || num_bands < 1 ||
        num_bands > I360HIP_MAX_BANDS)
        return I360HIP_INVALID_PARAM;
    const int L = ClampBands(width, height, num_bands);
    int w = width, h = height;
    const int K = L >= 4 ? L - 3 : 0;
    for (int l = 0; l < K; ++l) { w = (w + 1) / 2; h = (h + 1) / 2; }
    *level = K;
    *level_width = K ? w : 0;
    *level_height = K ? h : 0;
    return I360HIP_SUCCESS;
}

int i360hip_surrogate_features(i360hip_ctx* ctx, const i360hip_frame* frames, int num_frames,
                               const i360hip_chunk* chunks, int num_chunks, const int* contributors, int num_bands,
                               float* feat, float* target, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!feat || !target || (!frames && num_frames > 0) || (!chunks && num_chunks > 0) ||
        (!contributors && num_chunks > 0) || num_bands < 1 || num_bands > I360HIP_MAX_BANDS) {
        set_error(ctx, "Null pointer argument or numBands outside 1-12.");
        return I360HIP_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return I360HIP_DEVICE_ERROR;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    run_render(ctx, frames, num_frames, chunks, num_chunks, contributors, num_bands, I360HIP_RENDER_BATCHED, s, st,
               feat, target);
    if (!st.ok()) set_error(ctx, st.msg);
    return st.code;
}

int i360hip_surrogate_stats(i360hip_ctx* ctx, int* tiles, int* fallback_tiles) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!tiles || !fallback_tiles) return I360HIP_INVALID_PARAM;
    std::lock_guard<std::mutex> lk(ctx->mu);
    RenderCache& rc = cache_of(ctx);
    const Plan& P = rc.plan;
    if (P.surr != kSurrRender || P.stiles.empty()) {
        *tiles = 0;
        *fallback_tiles = 0;
        return I360HIP_SUCCESS;
    }
    std::vector<int> tf(P.stiles.size()), cf(P.chunks.size());
    if (hipSetDevice(ctx->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(tf.data(), P.sbufs.tflag, 4 * tf.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(cf.data(), P.sbufs.cflag, 4 * cf.size(), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error(ctx, "surrogate flag read-back failed");
        return I360HIP_DEVICE_ERROR;
    }
    int fb = 0;
    for (size_t t = 0; t < tf.size(); ++t) fb += tf[t] ? 1 : 0;
    *tiles = static_cast<int>(tf.size());
    *fallback_tiles = fb;
    return I360HIP_SUCCESS;
}

}  // extern "C"

extern "C" {

int i360hip_render_autotune(i360hip_ctx* ctx, const i360hip_frame* frames, int num_frames,
                            const i360hip_chunk* chunks, int num_chunks, const int* contributors, int num_bands,
                            i360hip_tuning* out) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!out || (!frames && num_frames > 0) || (!chunks && num_chunks > 0) || (!contributors && num_chunks > 0) ||
        num_bands < 1 || num_bands > I360HIP_MAX_BANDS) {
        set_error(ctx, "Null pointer argument or numBands outside 1-12.");
        return I360HIP_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return I360HIP_DEVICE_ERROR;
    hipStream_t s = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error(ctx, "autotune event creation failed");
        return I360HIP_DEVICE_ERROR;
    }
    // time one render (graph captured on the first call, then replayed) of
    // the batched form with the context's current configuration
    auto measure = [&](double& ms) {
        run_render(ctx, frames, num_frames, chunks, num_chunks, contributors, num_bands, I360HIP_RENDER_BATCHED, s,
                   st);  // capture (or eager) + first launch
        if (!st.ok() || hipStreamSynchronize(s) != hipSuccess) return false;
        float best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            (void)hipEventRecord(e0, s);
            run_render(ctx, frames, num_frames, chunks, num_chunks, contributors, num_bands, I360HIP_RENDER_BATCHED,
                       s, st);
            (void)hipEventRecord(e1, s);
            if (!st.ok() || hipEventSynchronize(e1) != hipSuccess) return false;
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, e0, e1);
            best = std::min(best, t);
        }
        ms = best;
        return true;
    };
    // coordinate search: chamfer strip, then chunk groups, then grid share
    int bestG = 4, bestS = 0, bestShare = 1;
    double bestMs = 1e300;
    auto trial = [&](int G, int S, int share) {
        ctx->tune_groups = G;
        ctx->tune_strip = S;
        ctx->tune_share = share;
        double ms = 0;
        if (!measure(ms)) return;
        if (ms < bestMs) { bestMs = ms; bestG = G; bestS = S; bestShare = share; }
    };
    for (int S : {4, 8, 16}) if (st.ok()) trial(4, S, 1);
    for (int G : {1, 2, 3}) if (st.ok()) trial(G, bestS, 1);
    if (st.ok() && bestG > 1) trial(bestG, bestS, 2);
    ctx->tune_groups = bestG;
    ctx->tune_strip = bestS;
    ctx->tune_share = bestShare;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!st.ok()) {
        set_error(ctx, st.msg);
        return st.code;
    }
    out->chunk_groups = bestG;
    out->chamfer_strip = bestS;
    out->grid_share = bestShare;
    out->ms = bestMs;
    return I360HIP_SUCCESS;
}

int i360hip_render_set_tuning(i360hip_ctx* ctx, const i360hip_tuning* tuning) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tune_groups = tuning ? std::max(0, tuning->chunk_groups) : 0;
    ctx->tune_strip = tuning ? std::max(0, tuning->chamfer_strip) : 0;
    ctx->tune_share = tuning ? std::max(0, tuning->grid_share) : 0;
    return I360HIP_SUCCESS;
}

}  // extern "C"
