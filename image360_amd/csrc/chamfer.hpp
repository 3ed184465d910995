// Coverage-distance (two-pass 3x3 chamfer) kernels shared by the RGB48 and
// float chunk renders.  Semantics: WindowsApp.Core/SourceFiles/
// SeamBlend.cpp:206-242, bit for bit.
//
// The serial raster loop makes pixel (x,y) of the forward pass depend on
// (x-1,y), (x,y-1), (x-1,y-1) and (x+1,y-1).  A block owns one
// (chunk, contributor) work item; thread t owns a strip of S columns and
// handles its r-th row at step 2r+t.  Every operand a pixel reads was
// therefore finished at an earlier step, and each pixel evaluates exactly
// the min() of the same binary32 sums the serial loop does (min is exact and
// order-free), so the result equals the serial one bit for bit.  The
// backward pass mirrors the schedule (thread order reversed).
//
// Only the contributor's coverage box is processed: pixels outside it are
// uncovered, and an uncovered pixel's distance is 0 in both passes, so the
// box rows/columns are surrounded by known zeros (or by the chunk edge,
// where the serial loop has no neighbour either).  Distances outside the
// box are never written; readers treat them as 0.
//
// Each thread prefetches its next P rows of input into registers, so the
// per-step critical path holds no global-memory latency.
#pragma once

#include "common.hpp"

namespace i360 {

struct ChamferWork {
    const uint8_t* cov;  // w*h coverage (0/1)
    float* dist;         // w*h distances, written inside the box only
    const int* box;      // x0, y0, x1, y1 (x1 < 0: nothing covered)
    int w, h;
};

constexpr float kChamferDiag = 1.41421356f;  // SeamBlend.cpp:209

template <int S>
struct StripIO;  // vector access to S consecutive values of one row
template <>
struct StripIO<4> {
    static __device__ __forceinline__ void cov(const uint8_t* p, float k, float (&d)[4]) {
        const uint32_t q = *reinterpret_cast<const uint32_t*>(p);
#pragma unroll
        for (int s = 0; s < 4; ++s) d[s] = ((q >> (8 * s)) & 0xffu) ? k : 0.0f;
    }
    static __device__ __forceinline__ void ld(const float* p, float (&d)[4]) {
        const float4 q = *reinterpret_cast<const float4*>(p);
        d[0] = q.x; d[1] = q.y; d[2] = q.z; d[3] = q.w;
    }
    static __device__ __forceinline__ void st(float* p, const float (&d)[4]) {
        *reinterpret_cast<float4*>(p) = make_float4(d[0], d[1], d[2], d[3]);
    }
};
template <int S>
struct StripIO {
    static __device__ __forceinline__ void cov(const uint8_t* p, float k, float (&d)[S]) {
#pragma unroll
        for (int q = 0; q < S / 4; ++q) StripIO<4>::cov(p + 4 * q, k, *reinterpret_cast<float(*)[4]>(d + 4 * q));
    }
    static __device__ __forceinline__ void ld(const float* p, float (&d)[S]) {
#pragma unroll
        for (int q = 0; q < S / 4; ++q) StripIO<4>::ld(p + 4 * q, *reinterpret_cast<float(*)[4]>(d + 4 * q));
    }
    static __device__ __forceinline__ void st(float* p, const float (&d)[S]) {
#pragma unroll
        for (int q = 0; q < S / 4; ++q) StripIO<4>::st(p + 4 * q, *reinterpret_cast<const float(*)[4]>(d + 4 * q));
    }
};

// Block layout: the first half of the wavefronts owns the even strips, the
// second half the odd ones.  Strips of one parity are busy on the same
// steps (row r of strip j runs at step 2r + j), so every wavefront is
// either fully busy or idle at each step.
template <int S, int P, bool FWD>
__global__ void __launch_bounds__(1024) chamfer_kernel(const ChamferWork* __restrict__ work) {
    __shared__ float firstv[2][1024];
    __shared__ float lastv[2][1024];
    const ChamferWork wk = work[blockIdx.x];
    const int x0 = wk.box[0], y0 = wk.box[1], x1 = wk.box[2], y1 = wk.box[3];
    if (x1 < 0) return;  // block-uniform
    const int w = wk.w, h = wk.h;
    const int xa = (x0 / S) * S;
    const int nact = (x1 + 1 - xa + S - 1) / S;
    const int rows = y1 - y0 + 1;
    const int half = blockDim.x >> 1;
    const int t = (threadIdx.x < half) ? 2 * threadIdx.x : 2 * (threadIdx.x - half) + 1;  // strip index
    const bool active = t < nact;
    const int xs = xa + t * S;
    const bool vec = (w & 3) == 0 && xs + S <= w;  // 16-byte aligned, whole strip inside the row
    const float kInf = static_cast<float>(w + h);
    const int lag = FWD ? t : (nact - 1 - t);
    const int base = lag >> 1, par = lag & 1;
    const bool lastT = (t == nact - 1);
    const bool edgePrev = FWD ? (y0 > 0) : (y1 < h - 1);  // a row exists beyond the box (all zeros)

    float ring[P][S];
    float prev[S];
#pragma unroll
    for (int s = 0; s < S; ++s) prev[s] = 0.0f;

    auto load_row = [&](int r, float (&dst)[S]) {
        const int y = FWD ? (y0 + r) : (y1 - r);
        const size_t rb = static_cast<size_t>(y) * w + xs;
        if (vec) {
            if (FWD) StripIO<S>::cov(wk.cov + rb, kInf, dst);
            else StripIO<S>::ld(wk.dist + rb, dst);
            return;
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bool in = xs + s < w;
            if (FWD) dst[s] = (in && wk.cov[rb + s]) ? kInf : 0.0f;
            else dst[s] = in ? wk.dist[rb + s] : 0.0f;
        }
    };

    // slot k serves rows k - base + j*P; fill it with its first row >= 0
#pragma unroll
    for (int k = 0; k < P; ++k) {
        int r = k - base;
        if (r < 0) r += ((-r + P - 1) / P) * P;
        if (active && r < rows) load_row(r, ring[k]);
    }

    const int tmax = 2 * (rows - 1) + nact - 1;
    for (int T0 = 0; T0 <= tmax; T0 += 2 * P) {
#pragma unroll
        for (int s2 = 0; s2 < 2 * P; ++s2) {
            if (T0 + s2 > tmax) break;  // uniform
            if (active && (s2 & 1) == par) {
                const int k = s2 >> 1;
                const int r = (T0 >> 1) + k - base;
                if (r >= 0 && r < rows) {
                    const int y = FWD ? (y0 + r) : (y1 - r);
                    const int pb = (r - 1) & 1, cb = r & 1;
                    const bool hasPrev = (r > 0) || edgePrev;
                    float cur[S];
                    if (FWD) {
                        const float carryL = (t > 0) ? lastv[cb][t - 1] : 0.0f;
                        const float prevL = (t > 0 && r > 0) ? lastv[pb][t - 1] : 0.0f;
                        const float prevR = (!lastT && r > 0) ? firstv[pb][t + 1] : 0.0f;
#pragma unroll
                        for (int s = 0; s < S; ++s) {
                            const int x = xs + s;
                            float b = ring[k][s];
                            if (x > 0) b = fminf(b, (s == 0 ? carryL : cur[s - 1]) + 1.0f);
                            if (hasPrev) {
                                b = fminf(b, prev[s] + 1.0f);
                                if (x > 0) b = fminf(b, (s == 0 ? prevL : prev[s - 1]) + kChamferDiag);
                                if (x < w - 1) b = fminf(b, (s == S - 1 ? prevR : prev[s + 1]) + kChamferDiag);
                            }
                            cur[s] = (x < w) ? b : 0.0f;
                        }
                    } else {
                        const float carryR = (!lastT) ? firstv[cb][t + 1] : 0.0f;
                        const float prevR = (!lastT && r > 0) ? firstv[pb][t + 1] : 0.0f;
                        const float prevL = (t > 0 && r > 0) ? lastv[pb][t - 1] : 0.0f;
#pragma unroll
                        for (int s = S - 1; s >= 0; --s) {
                            const int x = xs + s;
                            float b = ring[k][s];
                            if (x < w - 1) b = fminf(b, (s == S - 1 ? carryR : cur[s + 1]) + 1.0f);
                            if (hasPrev) {
                                b = fminf(b, prev[s] + 1.0f);
                                if (x < w - 1) b = fminf(b, (s == S - 1 ? prevR : prev[s + 1]) + kChamferDiag);
                                if (x > 0) b = fminf(b, (s == 0 ? prevL : prev[s - 1]) + kChamferDiag);
                            }
                            cur[s] = (x < w) ? b : 0.0f;
                        }
                    }
                    float* row = wk.dist + static_cast<size_t>(y) * w + xs;
                    if (vec) {
                        StripIO<S>::st(row, cur);
                    } else {
#pragma unroll
                        for (int s = 0; s < S; ++s)
                            if (xs + s < w) row[s] = cur[s];
                    }
#pragma unroll
                    for (int s = 0; s < S; ++s) prev[s] = cur[s];
                    firstv[cb][t] = cur[0];
                    lastv[cb][t] = cur[S - 1];
                    if (r + P < rows) load_row(r + P, ring[k]);
                }
            }
            lds_barrier();  // neighbours exchange through LDS only
        }
    }
}

// Threads per block for chunks of width w at strip width S.
inline int ChamferThreads(int w, int S) {
    const int cols = (w + S - 1) / S;
    return ((cols + 127) / 128) * 128;  // even and odd strips in separate wavefronts
}

inline int ChamferStrip(int w) { return w <= 4096 ? 4 : (w <= 8192 ? 8 : 16); }

// Launches forward then backward passes over `n` work items (device table).
inline hipError_t launch_chamfer(const ChamferWork* d_work, int n, int max_w, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int S = ChamferStrip(max_w);
    const int nt = ChamferThreads(max_w, S);
    if (S == 4) {
        chamfer_kernel<4, 8, true><<<n, nt, 0, s>>>(d_work);
        chamfer_kernel<4, 8, false><<<n, nt, 0, s>>>(d_work);
    } else if (S == 8) {
        chamfer_kernel<8, 4, true><<<n, nt, 0, s>>>(d_work);
        chamfer_kernel<8, 4, false><<<n, nt, 0, s>>>(d_work);
    } else {
        chamfer_kernel<16, 4, true><<<n, nt, 0, s>>>(d_work);
        chamfer_kernel<16, 4, false><<<n, nt, 0, s>>>(d_work);
    }
    return hipGetLastError();
}

}  // namespace i360
