// Bit-exact SeamBlend::BlendChunkContributors on the GPU
// (WindowsApp.Core/SourceFiles/SeamBlend.cpp:15-587).
//
// Stages, each a kernel family below:
//   1. coverage  : covered = any channel != 0, plus the coverage box
//                  (SeamBlend.cpp:560-568, 263-280).
//   2. chamfer   : the two-pass 3x3 chamfer (SeamBlend.cpp:206-242) with the
//                  exact float results of the serial raster order.  Pixel
//                  (x,y) of the forward pass depends on (x-1,y), (x,y-1),
//                  (x-1,y-1), (x+1,y-1); a block owns one contributor, each
//                  thread a strip of S columns, and thread t handles row r
//                  at step 2r+t (a systolic skew of two steps per strip), so
//                  every dependency was finished one step earlier and each
//                  pixel sees exactly the operands the serial loop sees.
//                  The backward pass mirrors it.
//   3. ownership : argmax distance > 0, ties to the lowest index
//                  (SeamBlend.cpp:244-261).
//   4. seam DP   : one block per contributor pair, in the reference's pair
//                  order (SeamBlend.cpp:573-584, 293-445): crossover and band
//                  per line with wavefront ballots, the line-to-line DP on
//                  one wavefront (lane = candidate position), backtrack
//                  through LDS-staged back-pointers, band reassignment.
//   5. pyramids  : per contributor, reduce levels 1..L-1 of the owner mask
//                  (weights) and of the hole-filled colour (numerator +
//                  denominator) with the separable [1,4,6,4,1]/16 blur
//                  evaluated only at the kept even samples
//                  (SeamBlend.cpp:15-79, 125-182).
//   6. compose   : per level, top-down, the weighted Laplacian sum over
//                  contributors in index order plus the expanded coarser
//                  reconstruction (SeamBlend.cpp:184-201, 501-525).  Level 0
//                  weights are the 0/1 owner mask, so only the owner's
//                  Laplacian reaches the level-0 band; zero-weight terms are
//                  skipped, which adds nothing (x + (+-0) == x).
// All float arithmetic keeps the reference's operation order; the library
// is compiled with -ffp-contract=off.

#include <climits>
#include <cstring>

#include "blend.hpp"
#include "chamfer.hpp"
#include "seam.hpp"
#include "ops.hpp"

namespace i360 {

void* ctx_scratch(i360hip_ctx* ctx, int slot, size_t bytes, Status& st);

namespace {

constexpr float kDiag = 1.41421356f;
constexpr int kSlotBlend = 8;

// ------------------------------------------------------------ coverage --
__global__ void __launch_bounds__(256) coverage_u16_kernel(PlaneTable imgs, int w, int h, uint8_t* __restrict__ cov,
                                                           int* __restrict__ boxes) {
    const int i = blockIdx.y;
    const size_t n = static_cast<size_t>(w) * h;
    const uint16_t* img = imgs.p[i];
    uint8_t* cv = cov + n * i;
    int bx0 = INT_MAX, by0 = INT_MAX, bx1 = -1, by1 = -1;
    for (size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < n;
         p += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const bool c = (img[p * 3] | img[p * 3 + 1] | img[p * 3 + 2]) != 0;
        cv[p] = c ? 1 : 0;
        if (c) {
            const int x = static_cast<int>(p % w), y = static_cast<int>(p / w);
            bx0 = min(bx0, x); by0 = min(by0, y); bx1 = max(bx1, x); by1 = max(by1, y);
        }
    }
    __shared__ int sb[4];
    if (threadIdx.x == 0) { sb[0] = INT_MAX; sb[1] = INT_MAX; sb[2] = -1; sb[3] = -1; }
    __syncthreads();
    if (bx1 >= 0) {
        atomicMin(&sb[0], bx0); atomicMin(&sb[1], by0); atomicMax(&sb[2], bx1); atomicMax(&sb[3], by1);
    }
    __syncthreads();
    if (threadIdx.x == 0 && sb[2] >= 0) {
        atomicMin(&boxes[4 * i + 0], sb[0]); atomicMin(&boxes[4 * i + 1], sb[1]);
        atomicMax(&boxes[4 * i + 2], sb[2]); atomicMax(&boxes[4 * i + 3], sb[3]);
    }
}

// Work table of the chamfer launch, written on the device (no host staging).
__global__ void chamfer_table_kernel(ChamferWork* work, const uint8_t* cov, uint8_t* scov, float* dist,
                                     const int* boxes, int c, int w, int h) {
    const int i = threadIdx.x;
    const size_t n = static_cast<size_t>(w) * h;
    const size_t np = static_cast<size_t>(DistPitch(w)) * ShearRows(w, h);
    if (i < c) work[i] = ChamferWork{cov + n * i, scov + np * i, dist + np * i, boxes + 4 * i, w, h};
}

__global__ void box_init_kernel(int* boxes, int c) {
    const int i = threadIdx.x;
    if (i < c) { boxes[4 * i] = INT_MAX; boxes[4 * i + 1] = INT_MAX; boxes[4 * i + 2] = -1; boxes[4 * i + 3] = -1; }
}

// ----------------------------------------------------------- ownership --
// Distances outside a contributor's coverage box are 0 (never written);
// the planes are the chamfer's sheared ones (shift = ChamferShear(w)).
__global__ void __launch_bounds__(256) ownership_kernel(const float* __restrict__ dist, const int* __restrict__ boxes,
                                                        int c, int w, size_t n, int shift, int32_t* __restrict__ owner) {
    const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int x = static_cast<int>(p % w), y = static_cast<int>(p / w);
    const size_t np = static_cast<size_t>(DistPitch(w)) * ShearRows(w, static_cast<int>(n / w));
    const size_t q = ShearAt(x, y, shift, DistPitch(w));
    int best = -1;
    float bd = 0.0f;
    for (int i = 0; i < c; ++i) {
        const int* b = boxes + 4 * i;
        if (x < b[0] || x > b[2] || y < b[1] || y > b[3]) continue;
        const float v = dist[np * i + q];
        if (v > bd) { bd = v; best = i; }
    }
    owner[p] = best;
}

// ------------------------------------------------------------ pyramids --
// Level l storage per contributor: a float4 plane (numerator rgb,
// denominator) and a float plane (weight).
constexpr int kRedTW = 64;   // outputs per block, x
constexpr int kRedTH = 8;    // outputs per block, y
constexpr int kRedRows = 2 * kRedTH + 3;

struct Level0Src {
    const uint16_t* img;
    const uint8_t* cov;
    const int32_t* owner;
    int self;
    __device__ __forceinline__ void load(size_t p, float (&v)[5]) const {
        const float m = cov[p] ? 1.0f : 0.0f;
        v[0] = static_cast<float>(img[p * 3 + 0]) * m;
        v[1] = static_cast<float>(img[p * 3 + 1]) * m;
        v[2] = static_cast<float>(img[p * 3 + 2]) * m;
        v[3] = m;
        v[4] = owner[p] == self ? 1.0f : 0.0f;
    }
};

struct LevelSrc {
    const float4* nd;
    const float* wt;
    __device__ __forceinline__ void load(size_t p, float (&v)[5]) const {
        const float4 q = nd[p];
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        v[4] = wt[p];
    }
};

struct PyrArgs {
    const uint16_t* img[I360HIP_MAX_CONTRIBUTORS];
    const uint8_t* cov;
    const int32_t* owner;
    const float4* nd_in[I360HIP_MAX_CONTRIBUTORS];
    const float* wt_in[I360HIP_MAX_CONTRIBUTORS];
    float4* nd_out[I360HIP_MAX_CONTRIBUTORS];
    float* wt_out[I360HIP_MAX_CONTRIBUTORS];
};

__device__ __forceinline__ float blur5(float a, float b, float m, float d, float e) {
    const float k0 = 1.0f / 16.0f, k1 = 4.0f / 16.0f, k2 = 6.0f / 16.0f;
    const float ta = k0 * a, tb = k1 * b, tm = k2 * m, td = k1 * d, te = k0 * e;
    return (((ta + tb) + tm) + td) + te;
}

// Reduce level l (w x h) of contributor blockIdx.z to level l+1.
template <bool kLevel0>
__global__ void __launch_bounds__(256) reduce_kernel(PyrArgs a, int w, int h, int ow, int oh) {
    __shared__ float tmp[5][kRedRows][kRedTW];
    const int ci = blockIdx.z;
    const int ox0 = blockIdx.x * kRedTW, oy0 = blockIdx.y * kRedTH;
    const size_t n = static_cast<size_t>(w) * h;
    // horizontal pass at the kept even columns for every needed row
    for (int idx = threadIdx.x; idx < kRedRows * kRedTW; idx += 256) {
        const int rr = idx / kRedTW, xx = idx % kRedTW;
        const int ox = ox0 + xx;
        if (ox >= ow) continue;
        const int sx = min(2 * ox, w - 1);
        const int ry = min(max(2 * oy0 - 2 + rr, 0), h - 1);
        const size_t rowb = static_cast<size_t>(ry) * w;
        float v[5][5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int cx = min(max(sx - 2 + k, 0), w - 1);
            if constexpr (kLevel0) {
                Level0Src src{a.img[ci], a.cov + n * ci, a.owner, ci};
                src.load(rowb + cx, v[k]);
            } else {
                LevelSrc src{a.nd_in[ci], a.wt_in[ci]};
                src.load(rowb + cx, v[k]);
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) tmp[c][rr][xx] = blur5(v[0][c], v[1][c], v[2][c], v[3][c], v[4][c]);
    }
    __syncthreads();
    // vertical pass at the kept even rows (rows clamp at the border exactly
    // as clampY does, since row ry above was clamped the same way)
    for (int idx = threadIdx.x; idx < kRedTH * kRedTW; idx += 256) {
        const int yy = idx / kRedTW, xx = idx % kRedTW;
        const int ox = ox0 + xx, oy = oy0 + yy;
        if (ox >= ow || oy >= oh) continue;
        const int r = 2 * yy + 2;  // row of sy = 2*oy in tmp
        float o[5];
#pragma unroll
        for (int c = 0; c < 5; ++c)
            o[c] = blur5(tmp[c][r - 2][xx], tmp[c][r - 1][xx], tmp[c][r][xx], tmp[c][r + 1][xx], tmp[c][r + 2][xx]);
        const size_t q = static_cast<size_t>(oy) * ow + ox;
        a.nd_out[ci][q] = make_float4(o[0], o[1], o[2], o[3]);
        a.wt_out[ci][q] = o[4];
    }
}

// Hole-filled colour at one sample: num * (den > 1e-6 ? 1/den : 0)
// (SeamBlend.cpp:177-179).
__device__ __forceinline__ float3 gauss_at(const float4* nd, size_t p) {
    const float4 q = nd[p];
    const float inv = (q.w > 1e-6f) ? __fdiv_rn(1.0f, q.w) : 0.0f;
    return make_float3(q.x * inv, q.y * inv, q.z * inv);
}

// Align-corners bilinear coordinates of SeamBlend.cpp:85-115.
struct ExpandCoord {
    int x0, x1, y0, y1;
    float wx, wy;
};

__device__ __forceinline__ ExpandCoord expand_coord(int x, int y, int sw, int sh, int dw, int dh) {
    const float sx = (sw > 1) ? __fdiv_rn(static_cast<float>(sw - 1), static_cast<float>(max(1, dw - 1))) : 0.0f;
    const float sy = (sh > 1) ? __fdiv_rn(static_cast<float>(sh - 1), static_cast<float>(max(1, dh - 1))) : 0.0f;
    ExpandCoord e;
    const float fy = sy * static_cast<float>(y);
    e.y0 = min(static_cast<int>(fy), sh - 1);
    e.y1 = min(e.y0 + 1, sh - 1);
    e.wy = fy - static_cast<float>(e.y0);
    const float fx = sx * static_cast<float>(x);
    e.x0 = min(static_cast<int>(fx), sw - 1);
    e.x1 = min(e.x0 + 1, sw - 1);
    e.wx = fx - static_cast<float>(e.x0);
    return e;
}

__device__ __forceinline__ float lerp2(const ExpandCoord& e, float v00, float v10, float v01, float v11) {
    const float omx = 1.0f - e.wx, omy = 1.0f - e.wy;
    const float top = omx * v00 + e.wx * v10;
    const float bot = omx * v01 + e.wx * v11;
    return omy * top + e.wy * bot;
}

__device__ __forceinline__ float3 expand_gauss(const float4* nd, int sw, const ExpandCoord& e) {
    const float3 a = gauss_at(nd, static_cast<size_t>(e.y0) * sw + e.x0);
    const float3 b = gauss_at(nd, static_cast<size_t>(e.y0) * sw + e.x1);
    const float3 c = gauss_at(nd, static_cast<size_t>(e.y1) * sw + e.x0);
    const float3 d = gauss_at(nd, static_cast<size_t>(e.y1) * sw + e.x1);
    return make_float3(lerp2(e, a.x, b.x, c.x, d.x), lerp2(e, a.y, b.y, c.y, d.y), lerp2(e, a.z, b.z, c.z, d.z));
}

__device__ __forceinline__ float3 expand_plain(const float4* rc, int sw, const ExpandCoord& e) {
    const float4 a = rc[static_cast<size_t>(e.y0) * sw + e.x0];
    const float4 b = rc[static_cast<size_t>(e.y0) * sw + e.x1];
    const float4 c = rc[static_cast<size_t>(e.y1) * sw + e.x0];
    const float4 d = rc[static_cast<size_t>(e.y1) * sw + e.x1];
    return make_float3(lerp2(e, a.x, b.x, c.x, d.x), lerp2(e, a.y, b.y, c.y, d.y), lerp2(e, a.z, b.z, c.z, d.z));
}

struct ComposeArgs {
    const float4* nd[I360HIP_MAX_CONTRIBUTORS];       // level l
    const float* wt[I360HIP_MAX_CONTRIBUTORS];        // level l
    const float4* nd_next[I360HIP_MAX_CONTRIBUTORS];  // level l+1 (null at the top)
    const float4* recon_next;                         // level l+1 reconstruction (null at the top)
    float4* recon;                                    // level l reconstruction
};

// recon_l = sum_i w_i * lap_i  (+ expand(recon_{l+1}) below the top).
__global__ void __launch_bounds__(256) compose_level_kernel(ComposeArgs a, int c, int w, int h, int nw, int nh,
                                                            int top) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t p = static_cast<size_t>(y) * w + x;
    float3 band = make_float3(0.0f, 0.0f, 0.0f);
    ExpandCoord e{};
    if (!top) e = expand_coord(x, y, nw, nh, w, h);
    for (int i = 0; i < c; ++i) {
        const float wt = a.wt[i][p];
        if (wt == 0.0f) continue;
        float3 lap = gauss_at(a.nd[i], p);
        if (!top) {
            const float3 up = expand_gauss(a.nd_next[i], nw, e);
            lap = make_float3(lap.x - up.x, lap.y - up.y, lap.z - up.z);
        }
        band.x = band.x + wt * lap.x;
        band.y = band.y + wt * lap.y;
        band.z = band.z + wt * lap.z;
    }
    if (!top) {
        const float3 up = expand_plain(a.recon_next, nw, e);
        band = make_float3(band.x + up.x, band.y + up.y, band.z + up.z);
    }
    a.recon[p] = make_float4(band.x, band.y, band.z, 0.0f);
}

struct FinalArgs {
    const uint16_t* img[I360HIP_MAX_CONTRIBUTORS];
    const uint8_t* cov;
    const float4* nd1[I360HIP_MAX_CONTRIBUTORS];  // level 1 of each contributor
    const float4* recon1;
    const int32_t* owner;
    uint16_t* out;
};

__device__ __forceinline__ unsigned short round_out(float v) {
    // (std::max)(0.0f, (std::min)(65535.0f, v)) then u16(v + 0.5f)
    // (SeamBlend.cpp:531-533)
    const float a = (v < 65535.0f) ? v : 65535.0f;
    const float b = (0.0f < a) ? a : 0.0f;
    return static_cast<unsigned short>(b + 0.5f);
}

// Level 0: only the owner's Laplacian carries weight (the 0/1 owner mask).
__global__ void __launch_bounds__(256) final_level0_kernel(FinalArgs a, int w, int h, int w1, int h1, int levels) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t p = static_cast<size_t>(y) * w + x;
    const size_t n = static_cast<size_t>(w) * h;
    const int o = a.owner[p];
    uint16_t* dst = a.out + p * 3;
    if (o < 0) { dst[0] = 0; dst[1] = 0; dst[2] = 0; return; }
    const float m = a.cov[n * o + p] ? 1.0f : 0.0f;
    const float inv = (m > 1e-6f) ? __fdiv_rn(1.0f, m) : 0.0f;
    const uint16_t* im = a.img[o] + p * 3;
    float g[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = (static_cast<float>(im[k]) * m) * inv;
    if (levels == 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) dst[k] = round_out(g[k]);
        return;
    }
    const ExpandCoord e = expand_coord(x, y, w1, h1, w, h);
    const float3 up = expand_gauss(a.nd1[o], w1, e);
    const float3 ur = expand_plain(a.recon1, w1, e);
    dst[0] = round_out((g[0] - up.x) + ur.x);
    dst[1] = round_out((g[1] - up.y) + ur.y);
    dst[2] = round_out((g[2] - up.z) + ur.z);
}

__global__ void __launch_bounds__(256) copy_owner_kernel(const int32_t* __restrict__ src, int32_t* __restrict__ dst,
                                                         size_t n) {
    const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[p];
}

__global__ void __launch_bounds__(256) zero_u16_kernel(uint16_t* __restrict__ d, size_t n) {
    const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) d[p] = 0;
}

__global__ void __launch_bounds__(256) fill_i32_kernel(int32_t* __restrict__ d, size_t n, int32_t v) {
    const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) d[p] = v;
}

inline size_t A256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

}  // namespace

void blend_release(i360hip_ctx* ctx) {
    if (static_cast<int>(ctx->bufs.size()) > kSlotBlend && ctx->bufs[kSlotBlend].ptr) {
        (void)hipFree(ctx->bufs[kSlotBlend].ptr);
        ctx->bufs[kSlotBlend] = DevBuf{};
    }
}

#define BL_LAUNCH(st, NAME, BYTES, ...)                                                        \
    do {                                                                                       \
        ProfScope ps_(ctx, s, NAME, static_cast<double>(BYTES));                               \
        __VA_ARGS__;                                                                           \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess) {                                                                \
            (st).set(I360HIP_KERNEL_LAUNCH_FAILED, "%s: %s", #__VA_ARGS__, HipName(e_));       \
            return;                                                                            \
        }                                                                                      \
    } while (0)

void blend_chunk_u16(i360hip_ctx* ctx, const uint16_t* const* warped, int c, int w, int h, int num_bands,
                     uint16_t* out, int32_t* owner_out, hipStream_t s, Status& st) {
    const size_t n = static_cast<size_t>(w) * h;
    if (c == 0) {
        BL_LAUNCH(st, "zero_u16", n * 6, zero_u16_kernel<<<GridFor(n * 3, 256), 256, 0, s>>>(out, n * 3));
        if (owner_out) BL_LAUNCH(st, "fill_i32", n * 4, fill_i32_kernel<<<GridFor(n, 256), 256, 0, s>>>(owner_out, n, -1));
        return;
    }
    if (w > 16384) {
        st.set(I360HIP_INVALID_PARAM, "chunk width %d exceeds the 16384-column limit of the chamfer kernel", w);
        return;
    }
    const int L = ClampBands(w, h, num_bands);
    int lw[I360HIP_MAX_BANDS], lh[I360HIP_MAX_BANDS];
    lw[0] = w;
    lh[0] = h;
    for (int l = 1; l < L; ++l) { lw[l] = (lw[l - 1] + 1) / 2; lh[l] = (lh[l - 1] + 1) / 2; }
    const int maxLines = w > h ? w : h;
    // workspace layout
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += A256(b); return o; };
    const size_t sheared = static_cast<size_t>(DistPitch(w)) * ShearRows(w, h);
    const size_t oCov = take(n * c), oScov = take(sheared * c), oDist = take(sheared * c * 4);
    const size_t oOwner = take(n * 4), oBox = take(4 * 4 * c);
    const size_t oLohi = take(8 * static_cast<size_t>(maxLines)), oPrev = take(4 * static_cast<size_t>(maxLines));
    const size_t oChosen = take(4 * static_cast<size_t>(maxLines)), oBack = take(64 * static_cast<size_t>(maxLines));
    const size_t oCost = take(64 * sizeof(float) * static_cast<size_t>(maxLines));
    const size_t oWork = take(sizeof(ChamferWork) * I360HIP_MAX_CONTRIBUTORS);
    const size_t oCnt = take(sizeof(int) * I360HIP_MAX_CONTRIBUTORS * I360HIP_MAX_CONTRIBUTORS);
    size_t oNd[I360HIP_MAX_BANDS][I360HIP_MAX_CONTRIBUTORS] = {}, oWt[I360HIP_MAX_BANDS][I360HIP_MAX_CONTRIBUTORS] = {};
    size_t oRec[I360HIP_MAX_BANDS] = {};
    for (int l = 1; l < L; ++l) {
        const size_t m = static_cast<size_t>(lw[l]) * lh[l];
        for (int i = 0; i < c; ++i) { oNd[l][i] = take(m * 16); oWt[l][i] = take(m * 4); }
        oRec[l] = take(m * 16);
    }
    auto* base = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotBlend, off, st));
    if (!st.ok()) return;
    auto* cov = base + oCov;
    auto* scov = base + oScov;
    auto* dist = reinterpret_cast<float*>(base + oDist);
    auto* owner = reinterpret_cast<int32_t*>(base + oOwner);
    auto* boxes = reinterpret_cast<int*>(base + oBox);
    PlaneTable imgs{};
    for (int i = 0; i < c; ++i) imgs.p[i] = warped[i];

    // 1. coverage + boxes
    BL_LAUNCH(st, "box_init", 16 * c, box_init_kernel<<<1, 64, 0, s>>>(boxes, c));
    {
        const unsigned gx = GridFor(n, 256) < 1024 ? GridFor(n, 256) : 1024;
        BL_LAUNCH(st, "coverage_u16", 7.0 * n * c, coverage_u16_kernel<<<dim3(gx, c), 256, 0, s>>>(imgs, w, h, cov, boxes));
    }
    // 2. chamfer, forward then backward (one block per contributor)
    {
        auto* dwork = reinterpret_cast<ChamferWork*>(base + oWork);
        BL_LAUNCH(st, "chamfer_table", 0, chamfer_table_kernel<<<1, 64, 0, s>>>(dwork, cov, scov, dist, boxes, c, w, h));
        {
            ProfScope ps(ctx, s, "chamfer", 13.0 * n * c);
            if (launch_chamfer(dwork, c, w, s) != hipSuccess) {
                st.set(I360HIP_KERNEL_LAUNCH_FAILED, "chamfer launch failed");
                return;
            }
        }
    }
    // 3. ownership
    BL_LAUNCH(st, "ownership", 4.0 * n * (c + 1), ownership_kernel<<<GridFor(n, 256), 256, 0, s>>>(dist, boxes, c, w, n,
                                                                                              ChamferShear(w), owner));
    // 4. pairwise seams in the reference's order
    int* counters = reinterpret_cast<int*>(base + oCnt);
    BL_LAUNCH(st, "fill_i32", 0, fill_i32_kernel<<<4, 256, 0, s>>>(counters, size_t(c) * c, 0));
    for (int i = 0; i < c; ++i)
        for (int j = i + 1; j < c; ++j) {
            int* lohi = reinterpret_cast<int*>(base + oLohi);
            int* chosen = reinterpret_cast<int*>(base + oChosen);
            int* cnt = counters + i * c + j;
            auto* costs = reinterpret_cast<float*>(base + oCost);
            const unsigned sb = (maxLines + kScanLinesPerBlock - 1) / kScanLinesPerBlock;
            BL_LAUNCH(st, "seam_scan", 0,
                      seam_scan_kernel<U16Img, float><<<sb, 256, 0, s>>>(U16Img{warped[i]}, U16Img{warped[j]},
                                                                         cov + n * i, cov + n * j, boxes, i, j, w,
                                                                         owner, lohi, chosen, cnt, costs));
            BL_LAUNCH(st, "seam_dp", 0, seam_dp_kernel<U16Img, float><<<1, kDpThreads, 0, s>>>(
                              boxes, i, j, costs, lohi, reinterpret_cast<int*>(base + oPrev), chosen, base + oBack, cnt));
            BL_LAUNCH(st, "seam_reassign", 0,
                      seam_reassign_kernel<><<<sb, 256, 0, s>>>(boxes, i, j, w, owner, lohi, chosen, cnt));
        }
    if (owner_out) BL_LAUNCH(st, "copy_owner", 8.0 * n, copy_owner_kernel<<<GridFor(n, 256), 256, 0, s>>>(owner, owner_out, n));
    // 5. pyramids of every contributor (contributors without owned pixels
    // carry all-zero weights and drop out of the sums exactly)
    PyrArgs pa{};
    for (int i = 0; i < c; ++i) pa.img[i] = warped[i];
    pa.cov = cov;
    pa.owner = owner;
    for (int l = 1; l < L; ++l) {
        for (int i = 0; i < c; ++i) {
            pa.nd_out[i] = reinterpret_cast<float4*>(base + oNd[l][i]);
            pa.wt_out[i] = reinterpret_cast<float*>(base + oWt[l][i]);
            if (l > 1) {
                pa.nd_in[i] = reinterpret_cast<const float4*>(base + oNd[l - 1][i]);
                pa.wt_in[i] = reinterpret_cast<const float*>(base + oWt[l - 1][i]);
            }
        }
        const dim3 grid((lw[l] + kRedTW - 1) / kRedTW, (lh[l] + kRedTH - 1) / kRedTH, c);
        if (l == 1) BL_LAUNCH(st, "reduce_l0", (11.0 * n + 20.0 * lw[1] * lh[1]) * c, reduce_kernel<true><<<grid, 256, 0, s>>>(pa, lw[0], lh[0], lw[1], lh[1]));
        else BL_LAUNCH(st, "reduce", 20.0 * c * (1.0 * lw[l - 1] * lh[l - 1] + 1.0 * lw[l] * lh[l]), reduce_kernel<false><<<grid, 256, 0, s>>>(pa, lw[l - 1], lh[l - 1], lw[l], lh[l]));
    }
    // 6. compose top-down, then level 0
    for (int l = L - 1; l >= 1; --l) {
        ComposeArgs ca{};
        for (int i = 0; i < c; ++i) {
            ca.nd[i] = reinterpret_cast<const float4*>(base + oNd[l][i]);
            ca.wt[i] = reinterpret_cast<const float*>(base + oWt[l][i]);
            if (l + 1 < L) ca.nd_next[i] = reinterpret_cast<const float4*>(base + oNd[l + 1][i]);
        }
        const int top = (l == L - 1) ? 1 : 0;
        ca.recon_next = top ? nullptr : reinterpret_cast<const float4*>(base + oRec[l + 1]);
        ca.recon = reinterpret_cast<float4*>(base + oRec[l]);
        const dim3 grid((lw[l] + 63) / 64, (lh[l] + 3) / 4);
        BL_LAUNCH(st, "compose_level", 1.0 * lw[l] * lh[l] * (20.0 * c + 16.0) + (top ? 0.0 : 1.0 * lw[l + 1] * lh[l + 1] * (16.0 * c + 16.0)), compose_level_kernel<<<grid, 256, 0, s>>>(ca, c, lw[l], lh[l], top ? 1 : lw[l + 1],
                                                               top ? 1 : lh[l + 1], top));
    }
    FinalArgs fa{};
    for (int i = 0; i < c; ++i) {
        fa.img[i] = warped[i];
        if (L > 1) fa.nd1[i] = reinterpret_cast<const float4*>(base + oNd[1][i]);
    }
    fa.cov = cov;
    fa.recon1 = L > 1 ? reinterpret_cast<const float4*>(base + oRec[1]) : nullptr;
    fa.owner = owner;
    fa.out = out;
    const dim3 grid((w + 63) / 64, (h + 3) / 4);
    BL_LAUNCH(st, "final_level0", 17.0 * n + (L > 1 ? 32.0 * lw[1] * lh[1] : 0.0), final_level0_kernel<<<grid, 256, 0, s>>>(fa, w, h, L > 1 ? lw[1] : 1, L > 1 ? lh[1] : 1, L));
}

}  // namespace i360
