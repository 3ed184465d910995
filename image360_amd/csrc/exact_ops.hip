// Bit-exact CDNA4 kernels for the IComputeBackend operations:
// median stack, Bayer demosaic, perspective warp (RGB48) and gain.
//
// Exactness discipline (the whole library is built with -ffp-contract=off):
// every binary32 operation below is a single correctly rounded operation in
// the order the reference Scalar tier evaluates it; division and square
// root use the correctly rounded forms (__fdiv_rn, __fsqrt_rn); float->u16
// conversions truncate after the reference's [0, 65535] clamp.

#include "ops.hpp"

namespace i360 {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ us2 mk2(unsigned short a, unsigned short b) {
    us2 r;
    r.x = a;
    r.y = b;
    return r;
}

__device__ __forceinline__ unsigned short sat_u16(float v) {
    // (std::min)((std::max)(v, 0.0f), 65535.0f), then truncation
    float a = (v < 0.0f) ? 0.0f : v;
    float b = (65535.0f < a) ? 65535.0f : a;
    return static_cast<unsigned short>(b);
}

// ======================================================================
// Median stack (MedianStackKernels.h:19-55, MedianStack_Scalar.cpp:8-24).
//
// One thread owns 2*PAIRS consecutive scalars and keeps its NP values per
// scalar in packed u16 pairs, so the sorting network's compare-exchanges
// are v_pk_min_u16 / v_pk_max_u16 on two scalars at once.  Inputs beyond
// the real count, and values the sigma clip rejects, are replaced by
// 0xFFFF: after sorting, the kept values occupy the front in order, so the
// upper median of the kept set is element kept/2 (a kept 0xFFFF equals the
// sentinel, which leaves the answer unchanged).
// ======================================================================

template <int NP>
__device__ __forceinline__ void bitonic_sort(us2 (&v)[NP]) {
#pragma unroll
    for (int k = 2; k <= NP; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    us2 a = v[i], b = v[l];
                    us2 lo = __builtin_elementwise_min(a, b);
                    us2 hi = __builtin_elementwise_max(a, b);
                    if ((i & k) == 0) { v[i] = lo; v[l] = hi; }
                    else { v[i] = hi; v[l] = lo; }
                }
            }
        }
    }
}

// Sigma-clip one scalar (component `C` of the packed registers); returns
// the index of the upper median after sorting.
template <int NP, int C>
__device__ __forceinline__ int clip_component(us2 (&v)[NP], int count, float thr) {
    if (count <= 3) return count / 2;
    float sum = 0.0f, sq = 0.0f;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        if (i < count) {
            float f = static_cast<float>(v[i][C]);
            sum = sum + f;
            sq = sq + f * f;
        }
    }
    const float cnt = static_cast<float>(count);
    const float mean = __fdiv_rn(sum, cnt);
    const float var = __fdiv_rn(sq, cnt) - mean * mean;
    const float sigma = __fsqrt_rn(var > 0.0f ? var : 0.0f);
    const float ts = thr * sigma;
    const float lower = mean - ts;
    const float upper = mean + ts;
    int kept = 0;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        if (i < count) {
            float f = static_cast<float>(v[i][C]);
            kept += (f >= lower && f <= upper) ? 1 : 0;
        }
    }
    if (kept == 0) return count / 2;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        if (i < count) {
            float f = static_cast<float>(v[i][C]);
            if (!(f >= lower && f <= upper)) v[i][C] = 0xFFFF;
        }
    }
    return kept / 2;
}

template <int NP>
__device__ __forceinline__ void median_pair(us2 (&v)[NP], int count, float thr, unsigned short& o0,
                                            unsigned short& o1) {
    const int k0 = clip_component<NP, 0>(v, count, thr);
    const int k1 = clip_component<NP, 1>(v, count, thr);
    bitonic_sort<NP>(v);
    unsigned short r0 = v[0][0], r1 = v[0][1];
#pragma unroll
    for (int i = 1; i < NP; ++i) {
        r0 = (i == k0) ? v[i][0] : r0;
        r1 = (i == k1) ? v[i][1] : r1;
    }
    o0 = r0;
    o1 = r1;
}

// Each thread: 4 consecutive scalars (one 8-byte load per plane) when the
// planes are 8-byte aligned; the tail and unaligned planes take the
// 2-scalar path with 4-byte loads, or single scalars at the very end.
template <int NP>
__global__ void __launch_bounds__(256) median_stack_kernel(PlaneTable planes, int count, uint16_t* __restrict__ out,
                                                           size_t n, float thr, int vec4) {
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (vec4) {
        const size_t base = t * 4;
        if (base + 4 <= n) {
            us2 a[NP], b[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                if (i < count) {
                    uint2 q = *reinterpret_cast<const uint2*>(planes.p[i] + base);
                    a[i] = __builtin_bit_cast(us2, q.x);
                    b[i] = __builtin_bit_cast(us2, q.y);
                } else {
                    a[i] = mk2(0xFFFF, 0xFFFF);
                    b[i] = mk2(0xFFFF, 0xFFFF);
                }
            }
            unsigned short o0, o1, o2, o3;
            median_pair<NP>(a, count, thr, o0, o1);
            median_pair<NP>(b, count, thr, o2, o3);
            uint2 r;
            r.x = static_cast<uint32_t>(o0) | (static_cast<uint32_t>(o1) << 16);
            r.y = static_cast<uint32_t>(o2) | (static_cast<uint32_t>(o3) << 16);
            *reinterpret_cast<uint2*>(out + base) = r;
            return;
        }
        // tail of a vectorised launch: fall through to per-scalar handling
        for (size_t s = base; s < n && s < base + 4; ++s) {
            us2 a[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                unsigned short x = (i < count) ? planes.p[i][s] : 0xFFFF;
                a[i] = mk2(x, 0xFFFF);
            }
            unsigned short o0, o1;
            median_pair<NP>(a, count, thr, o0, o1);
            out[s] = o0;
        }
        return;
    }
    // unaligned planes: two scalars per thread, element loads
    const size_t base = t * 2;
    if (base >= n) return;
    const bool two = base + 1 < n;
    us2 a[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        if (i < count) {
            unsigned short x0 = planes.p[i][base];
            unsigned short x1 = two ? planes.p[i][base + 1] : static_cast<unsigned short>(0xFFFF);
            a[i] = mk2(x0, x1);
        } else {
            a[i] = mk2(0xFFFF, 0xFFFF);
        }
    }
    unsigned short o0, o1;
    median_pair<NP>(a, two ? count : count, thr, o0, o1);
    out[base] = o0;
    if (two) out[base + 1] = o1;
}

hipError_t launch_median_stack(const PlaneTable& planes, int count, uint16_t* out, size_t n, float thr,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    bool aligned = (reinterpret_cast<uintptr_t>(out) & 7u) == 0;
    for (int i = 0; i < count; ++i) aligned = aligned && ((reinterpret_cast<uintptr_t>(planes.p[i]) & 7u) == 0);
    const int vec4 = aligned ? 1 : 0;
    const size_t per_thread = aligned ? 4 : 2;
    const dim3 block(256);
    const dim3 grid(GridFor(n, 256 * per_thread));
    if (count <= 4) median_stack_kernel<4><<<grid, block, 0, s>>>(planes, count, out, n, thr, vec4);
    else if (count <= 8) median_stack_kernel<8><<<grid, block, 0, s>>>(planes, count, out, n, thr, vec4);
    else if (count <= 16) median_stack_kernel<16><<<grid, block, 0, s>>>(planes, count, out, n, thr, vec4);
    else median_stack_kernel<32><<<grid, block, 0, s>>>(planes, count, out, n, thr, vec4);
    return hipGetLastError();
}

// ======================================================================
// Bayer demosaic (BayerDemosaic_Scalar.cpp:9-50, BayerDemosaicKernels.h).
//
// A 256-thread block covers a 64x16 pixel tile.  The (64+2)x(16+2) CFA halo
// tile is loaded once, black-level-subtracted and white-balanced (u16, as
// the reference stores it) into LDS; every thread then forms 4 output
// pixels from LDS.  Same-colour sums are exact integers in binary32 (at most
// 4 * 65535 < 2^24), so only the division and the colour matrix need the
// reference's operation order.
// ======================================================================

constexpr int kDmTileW = 64;
constexpr int kDmTileH = 16;

__device__ __forceinline__ int cfa_color(int row, int col, uint32_t filters) {
    return static_cast<int>((filters >> ((((row << 1) & 14) | (col & 1)) << 1)) & 3u);
}

template <bool kFloatOut>
__global__ void __launch_bounds__(256) demosaic_kernel(const uint16_t* __restrict__ cfa, int w, int h,
                                                       unsigned short black, DemosaicParams prm,
                                                       void* __restrict__ out) {
    __shared__ unsigned short tile[kDmTileH + 2][kDmTileW + 2];
    const int tx0 = blockIdx.x * kDmTileW;
    const int ty0 = blockIdx.y * kDmTileH;
    const int tid = threadIdx.x;
    for (int i = tid; i < (kDmTileW + 2) * (kDmTileH + 2); i += 256) {
        const int ly = i / (kDmTileW + 2), lx = i % (kDmTileW + 2);
        const int gx = tx0 + lx - 1, gy = ty0 + ly - 1;
        unsigned short v = 0;
        if (gx >= 0 && gy >= 0 && gx < w && gy < h) {
            unsigned short raw = cfa[static_cast<size_t>(gy) * w + gx];
            unsigned short bl = raw > black ? static_cast<unsigned short>(raw - black) : 0;
            v = sat_u16(static_cast<float>(bl) * prm.cam_mul[cfa_color(gy, gx, prm.filters)]);
        }
        tile[ly][lx] = v;
    }
    __syncthreads();
    // 4 pixels per thread: rows ty0 + (tid/64)*4 .. +3, column tx0 + tid%64
    const int lx = tid & 63;
    const int x = tx0 + lx;
    if (x >= w) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ly = (tid >> 6) * 4 + r;
        const int y = ty0 + ly;
        if (y >= h) break;
        float acc[3] = {0.0f, 0.0f, 0.0f};
        int cnt[3] = {0, 0, 0};
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = x + dx, ny = y + dy;
                if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
                const int ch = cfa_color(ny, nx, prm.filters);
                const int slot = ch == 3 ? 1 : ch;
                const float v = static_cast<float>(tile[ly + 1 + dy][lx + 1 + dx]);
                acc[0] = slot == 0 ? acc[0] + v : acc[0];
                acc[1] = slot == 1 ? acc[1] + v : acc[1];
                acc[2] = slot == 2 ? acc[2] + v : acc[2];
                cnt[0] += slot == 0;
                cnt[1] += slot == 1;
                cnt[2] += slot == 2;
            }
        }
        const int mine = cfa_color(y, x, prm.filters);
        const int ms = mine == 3 ? 1 : mine;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (c != ms && cnt[c] != 0) acc[c] = __fdiv_rn(acc[c], static_cast<float>(cnt[c]));
        const float own = static_cast<float>(tile[ly + 1][lx + 1]);
        acc[0] = ms == 0 ? own : acc[0];
        acc[1] = ms == 1 ? own : acc[1];
        acc[2] = ms == 2 ? own : acc[2];
        const float rr = static_cast<float>(sat_u16(acc[0]));
        const float gg = static_cast<float>(sat_u16(acc[1]));
        const float bb = static_cast<float>(sat_u16(acc[2]));
        const size_t o = (static_cast<size_t>(y) * w + x) * 3;
#pragma unroll
        for (int row = 0; row < 3; ++row) {
            const float t0 = prm.rgb_cam[row * 4 + 0] * rr;
            const float t1 = prm.rgb_cam[row * 4 + 1] * gg;
            const float t2 = prm.rgb_cam[row * 4 + 2] * bb;
            const float val = ((t0 + t1) + t2) + prm.rgb_cam[row * 4 + 3];
            if constexpr (kFloatOut) {
                const float a = (val < 0.0f) ? 0.0f : val;
                static_cast<float*>(out)[o + row] = (65535.0f < a) ? 65535.0f : a;
            } else {
                static_cast<uint16_t*>(out)[o + row] = sat_u16(val);
            }
        }
    }
}

hipError_t launch_demosaic(const uint16_t* cfa, int w, int h, unsigned short black, const DemosaicParams& prm,
                           void* out, bool float_out, hipStream_t s) {
    const dim3 block(256);
    const dim3 grid((w + kDmTileW - 1) / kDmTileW, (h + kDmTileH - 1) / kDmTileH);
    if (float_out) demosaic_kernel<true><<<grid, block, 0, s>>>(cfa, w, h, black, prm, out);
    else demosaic_kernel<false><<<grid, block, 0, s>>>(cfa, w, h, black, prm, out);
    return hipGetLastError();
}

// ======================================================================
// Perspective warp, RGB48 (WarpPerspective_Scalar.cpp:8-45 and
// WarpPerspectiveKernels.h:16-43).  One thread per destination pixel.
// ======================================================================

__global__ void __launch_bounds__(256) warp_u16_kernel(const uint16_t* __restrict__ src, int sw, int sh,
                                                       uint16_t* __restrict__ dst, int dw, int dh, Mat3f m,
                                                       int ox, int oy) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= dw || y >= dh) return;
    const float X = static_cast<float>(x + ox);
    const float Y = static_cast<float>(y + oy);
    float u = (m.v[0] * X + m.v[1] * Y) + m.v[2];
    float v = (m.v[3] * X + m.v[4] * Y) + m.v[5];
    const float q = (m.v[6] * X + m.v[7] * Y) + m.v[8];
    uint16_t* o = dst + (static_cast<size_t>(y) * dw + x) * 3;
    if (fabsf(q) < 1e-10f) { o[0] = 0; o[1] = 0; o[2] = 0; return; }
    u = __fdiv_rn(u, q);
    v = __fdiv_rn(v, q);
    const float fu = floorf(u), fv = floorf(v);
    // sx0 >= 0 && sy0 >= 0 && sx0+1 < srcW && sy0+1 < srcH, decided on the
    // floored float so out-of-range and NaN coordinates never reach an
    // int conversion (the x86 conversion the reference relies on yields
    // INT_MIN for them, which its test also rejects).
    if (!(fu >= 0.0f && fv >= 0.0f && fu <= static_cast<float>(sw - 2) && fv <= static_cast<float>(sh - 2))) {
        o[0] = 0; o[1] = 0; o[2] = 0;
        return;
    }
    const int sx0 = static_cast<int>(fu), sy0 = static_cast<int>(fv);
    const float fx = u - static_cast<float>(sx0);
    const float fy = v - static_cast<float>(sy0);
    const float omx = 1.0f - fx, omy = 1.0f - fy;
    const float w00 = omx * omy, w10 = fx * omy, w01 = omx * fy, w11 = fx * fy;
    const uint16_t* r0 = src + (static_cast<size_t>(sy0) * sw + sx0) * 3;
    const uint16_t* r1 = r0 + static_cast<size_t>(sw) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float a = w00 * static_cast<float>(r0[c]);
        const float b = w10 * static_cast<float>(r0[3 + c]);
        const float d = w01 * static_cast<float>(r1[c]);
        const float e = w11 * static_cast<float>(r1[3 + c]);
        o[c] = sat_u16(((a + b) + d) + e);
    }
}

hipError_t launch_warp_u16(const uint16_t* src, int sw, int sh, uint16_t* dst, int dw, int dh, const Mat3f& m,
                           int ox, int oy, hipStream_t s) {
    const dim3 block(256);
    const dim3 grid((dw + 63) / 64, (dh + 3) / 4);
    warp_u16_kernel<<<grid, block, 0, s>>>(src, sw, sh, dst, dw, dh, m, ox, oy);
    return hipGetLastError();
}

// ======================================================================
// Gain (GainColorOps.cpp:16-26): eight u16 per thread (one 16-byte access)
// when aligned.
// ======================================================================

__global__ void __launch_bounds__(256) gain_kernel(uint16_t* __restrict__ d, size_t n, float g, int vec8) {
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (vec8) {
        const size_t base = t * 8;
        if (base + 8 <= n) {
            uint4 q = *reinterpret_cast<const uint4*>(d + base);
            uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const unsigned short a = sat_u16(static_cast<float>(w[k] & 0xFFFFu) * g);
                const unsigned short b = sat_u16(static_cast<float>(w[k] >> 16) * g);
                w[k] = static_cast<uint32_t>(a) | (static_cast<uint32_t>(b) << 16);
            }
            *reinterpret_cast<uint4*>(d + base) = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
            for (size_t i = base; i < n; ++i) d[i] = sat_u16(static_cast<float>(d[i]) * g);
        }
        return;
    }
    if (t < n) d[t] = sat_u16(static_cast<float>(d[t]) * g);
}

hipError_t launch_gain_u16(uint16_t* d, size_t n, float g, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int vec8 = (reinterpret_cast<uintptr_t>(d) & 15u) == 0;
    const dim3 block(256);
    const dim3 grid(GridFor(n, vec8 ? 256 * 8 : 256));
    gain_kernel<<<grid, block, 0, s>>>(d, n, g, vec8);
    return hipGetLastError();
}

}  // namespace i360
