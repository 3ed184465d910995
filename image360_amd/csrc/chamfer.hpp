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
//
// Sheared planes: at a given step, strip t works on row y0 + (T - t)/2, so
// in a row-major plane the 64 lanes of a wavefront would touch 64 different
// rows.  Distances (and a copy of the coverage the forward pass reads) are
// therefore stored sheared: pixel (x, y) lives at row y + ((x / S) >> 1) of
// its plane.  All strips of one parity, one wavefront, then touch a single
// row at every step, in both passes, and their loads and stores coalesce.
#pragma once

#include <cstdlib>

#include "common.hpp"

namespace i360 {

struct ChamferWork {
    const uint8_t* cov;  // w*h coverage (0/1), row pitch w
    uint8_t* scov;       // sheared copy of the coverage box (written by chamfer_shear_kernel)
    float* dist;         // sheared distances, written inside the box only
    const int* box;      // x0, y0, x1, y1 (x1 < 0: nothing covered)
    int w, h;
};

// Row pitch (elements) of the sheared planes: a multiple of 16, so a strip
// of 4, 8 or 16 columns starting at a multiple of its width never crosses
// its row.
__host__ __device__ inline int DistPitch(int w) { return (w + 15) & ~15; }

// Narrowest strip a launch may pick for chunks of width w (one strip per
// thread, at most 1024 threads).
__host__ __device__ inline int ChamferMinStrip(int w) { return w <= 4096 ? 4 : (w <= 8192 ? 8 : 16); }

// Rows of a sheared plane of an h-row chunk: room for the shear of the
// last column at the narrowest strip.
__host__ __device__ inline int ShearRows(int w, int h) { return h + (((w - 1) / ChamferMinStrip(w)) >> 1) + 1; }

// Element offset of pixel (x, y) in a sheared plane; shift = log2(S) + 1.
__host__ __device__ inline size_t ShearAt(int x, int y, int shift, int pitch) {
    return static_cast<size_t>(y + (x >> shift)) * pitch + x;
}
__host__ __device__ constexpr int ShearShift(int S) { return S == 4 ? 3 : (S == 8 ? 4 : 5); }

constexpr float kChamferDiag = 1.41421356f;  // SeamBlend.cpp:209

typedef float f4v __attribute__((ext_vector_type(4)));
typedef uint32_t u2v __attribute__((ext_vector_type(2)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t ld_u8(const uint8_t* p) { return *gptr(p); }
__device__ __forceinline__ f4v ld_f4(const float* p) { return *gptr(reinterpret_cast<const f4v*>(p)); }
__device__ __forceinline__ void st_f4(float* p, f4v v) { *gptr(reinterpret_cast<f4v*>(p)) = v; }

// Copies each item's coverage box, widened to whole 16-column groups, into
// its sheared plane, 8 bytes per thread (an 8-column group never straddles
// two shear rows for S >= 4); columns at or past w are written as 0.
// Grid: (blocks per item, items).
template <int S>
__global__ void __launch_bounds__(256) chamfer_shear_kernel(const ChamferWork* __restrict__ work) {
    const ChamferWork wk = work[blockIdx.y];
    const int x0 = wk.box[0], y0 = wk.box[1], x1 = wk.box[2], y1 = wk.box[3];
    if (x1 < 0) return;
    const int w = wk.w, cp = DistPitch(w);
    const int ga = x0 & ~15, gb = min((x1 + 16) & ~15, cp);
    const int ng = (gb - ga) >> 3;
    const size_t total = static_cast<size_t>(ng) * (y1 - y0 + 1);
    const bool aligned = (w & 7) == 0;
    for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int y = y0 + static_cast<int>(e / ng), x = ga + 8 * static_cast<int>(e % ng);
        const uint8_t* c = wk.cov + static_cast<size_t>(y) * w + x;
        uint32_t lo = 0, hi = 0;
        if (aligned && x + 8 <= w) {
            const u2v q = *gptr(reinterpret_cast<const u2v*>(c));
            lo = q.x; hi = q.y;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (x + k < w) lo |= ld_u8(c + k) << (8 * k);
                if (x + 4 + k < w) hi |= ld_u8(c + 4 + k) << (8 * k);
            }
        }
        u2v v;
        v.x = lo; v.y = hi;
        *gptr(reinterpret_cast<u2v*>(wk.scov + ShearAt(x, y, ShearShift(S), cp))) = v;
    }
}

// Block layout: the first half of the wavefronts owns the even strips, the
// second half the odd ones.  Strips of one parity are busy on the same
// steps (row r of strip j runs at step 2r + j), so every wavefront is
// either fully busy or idle at each step.
//
// Neighbours beyond the processed area are constants: a pixel that exists
// outside the coverage box is uncovered (0), a pixel beyond the chunk edge
// does not exist (+inf, which no min picks).  They come from sentinel LDS
// slots (left of the first strip, right of the last) and, for the row
// before the box, from per-thread constants, so every LDS read is
// unconditional and a pixel's update is branch-free.  Distances are
// non-negative floats, whose order is their bit patterns' unsigned order:
// min is one integer instruction.
__device__ __forceinline__ uint32_t dmin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__host__ __device__ constexpr int RWmax(int S) { return S / 4; }
__device__ __forceinline__ uint32_t dadd(uint32_t a, float c) { return __float_as_uint(__uint_as_float(a) + c); }

template <int S, int P, bool FWD>
__global__ void __launch_bounds__(1024) chamfer_kernel(const ChamferWork* __restrict__ work) {
    // boundary columns of every strip for the last two rows: slot t + 1
    // holds strip t; slots 0 and nact + 1 the sentinels
    __shared__ uint32_t firstv[2][1024 + 2];
    __shared__ uint32_t lastv[2][1024 + 2];
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
    const int dp = DistPitch(w);
    const int xl = xa + min(t, nact - 1) * S;  // load strip: inactive threads load the last active one
    constexpr int SH = ShearShift(S);
    constexpr uint32_t kNone = 0x7f800000u;  // +inf: no pixel there
    const float kInf = static_cast<float>(w + h);
    const int lag = FWD ? t : (nact - 1 - t);
    const int base = lag >> 1, par = lag & 1;
    const bool edgePrev = FWD ? (y0 > 0) : (y1 < h - 1);  // a row exists before the box (all uncovered)

    // per-column constants of this strip
    uint32_t inBox = 0, exists = 0;  // bit s: column xs + s
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int x = xs + s;
        inBox |= (x >= x0 && x <= x1) ? (1u << s) : 0u;
        exists |= (x < w) ? (1u << s) : 0u;
    }
    // FWD: coverage bytes of the columns outside the box are masked to 0
    uint32_t byteMask[RWmax(S)];
#pragma unroll
    for (int g = 0; g < S / 4; ++g) {
        uint32_t m = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) m |= ((inBox >> (4 * g + b)) & 1u) ? (0xffu << (8 * b)) : 0u;
        byteMask[g] = m;
    }
    // the row before the box, and the columns left and right of the strip there
    const uint32_t pL0 = (edgePrev && xs > 0) ? 0u : kNone;
    const uint32_t pR0 = (edgePrev && xs + S < w) ? 0u : kNone;
    if (threadIdx.x == 0) {
        const uint32_t sl = xa > 0 ? 0u : kNone, sr = (xa + nact * S < w) ? 0u : kNone;
        lastv[0][0] = sl; lastv[1][0] = sl;
        firstv[0][nact + 1] = sr; firstv[1][nact + 1] = sr;
    }

    // raw prefetched row data; converted only when the row is processed so
    // that no instruction depends on a load until its row's step
    //   FWD: the strip's S coverage bytes, packed four to a word
    //   BWD: the forward distance bits
    // Three row buffers rotate (no register copies): superstep j reads
    // buffer j mod 3 and issues the loads of superstep j + 2.  A copy, or a
    // load one superstep ahead, would make the compiler wait for every
    // older memory operation (loads and the conditional stores share one
    // in-order counter) at each superstep; two ahead, the wait at a
    // buffer's first use leaves the newest loads and stores in flight.
    constexpr int RW = FWD ? S / 4 : S;
    uint32_t bufA[P][RW], bufB[P][RW], bufC[P][RW];
    uint32_t prev[S];
#pragma unroll
    for (int s = 0; s < S; ++s) prev[s] = (edgePrev && ((exists >> s) & 1u)) ? 0u : kNone;

    // loads are unconditional (rows and strips clamped into the box) so
    // that no register is merged across a branch while its load is in flight
    auto load_row = [&](int r, uint32_t (&dst)[RW]) {
        const int y = FWD ? (y0 + r) : (y1 - r);
        if constexpr (FWD) {
            const uint8_t* c = wk.scov + ShearAt(xl, y, SH, dp);
            if constexpr (S == 4) {
                dst[0] = *gptr(reinterpret_cast<const uint32_t*>(c));
            } else if constexpr (S == 8) {
                const u2v q = *gptr(reinterpret_cast<const u2v*>(c));
                dst[0] = q.x; dst[1] = q.y;
            } else {
                const u4v q = *gptr(reinterpret_cast<const u4v*>(c));
                dst[0] = q.x; dst[1] = q.y; dst[2] = q.z; dst[3] = q.w;
            }
        } else {
            const float* d = wk.dist + ShearAt(xl, y, SH, dp);
#pragma unroll
            for (int g = 0; g < S / 4; ++g) {
                const f4v q = ld_f4(d + 4 * g);
                dst[4 * g + 0] = __float_as_uint(q.x);
                dst[4 * g + 1] = __float_as_uint(q.y);
                dst[4 * g + 2] = __float_as_uint(q.z);
                dst[4 * g + 3] = __float_as_uint(q.w);
            }
        }
    };
    // initial value of column s: covered -> w + h (FWD) / the forward
    // distance (BWD); outside the box 0
    auto init_of = [&](const uint32_t (&src)[RW], int s) -> uint32_t {
        if constexpr (FWD) {
            // coverage byte (0 or 1; 0 outside the box) times w + h
            const float c = static_cast<float>(((src[s >> 2] & byteMask[s >> 2]) >> (8 * (s & 3))) & 0xffu);
            return __float_as_uint(c * kInf);
        } else {
            return ((inBox >> s) & 1u) ? src[s] : 0u;
        }
    };
    // superstep j (steps 2Pj .. 2Pj+2P-1) processes rows Pj - base + k, k < P,
    // in slot k of its buffer
    auto fetch = [&](int j, uint32_t (&dst)[P][RW]) {
#pragma unroll
        for (int k = 0; k < P; ++k) load_row(min(max(P * j + k - base, 0), rows - 1), dst[k]);
    };
    fetch(0, bufA);
    fetch(1, bufB);
    // row r of this strip lives at pbase +/- r * dp (sheared plane; r * dp < 2^28)
    float* const pbase = wk.dist + ShearAt(xs, FWD ? y0 : y1, SH, dp);
    lds_barrier();  // sentinels visible

    const int tmax = 2 * (rows - 1) + nact - 1;
    // one superstep; false (block-uniform) once past the last step
    auto superstep = [&](int j, const uint32_t (&ring)[P][RW], uint32_t (&ld)[P][RW]) -> bool {
        const int T0 = 2 * P * j;
        if (T0 > tmax) return false;
        fetch(j + 2, ld);
        const int nsteps = min(2 * P, tmax - T0 + 1);  // uniform
#pragma unroll
        for (int s2 = 0; s2 < 2 * P; ++s2) {
            if (s2 >= nsteps) continue;  // uniform: all threads skip the barrier together
            if (active && (s2 & 1) == par) {
                const int k = s2 >> 1;
                const int r = (T0 >> 1) + k - base;
                if (r >= 0 && r < rows) {
                    const int pb = (r - 1) & 1, cb = r & 1;
                    // neighbours (slot = strip + 1), all read together
                    const uint32_t lCur = lastv[cb][t], rCur = firstv[cb][t + 2];
                    const uint32_t lPrev = lastv[pb][t], rPrev = firstv[pb][t + 2];
                    const bool first = r == 0;
                    const uint32_t prevL = first ? pL0 : lPrev;
                    const uint32_t prevR = first ? pR0 : rPrev;
                    uint32_t cur[S];
                    if constexpr (FWD) {
#pragma unroll
                        for (int s = 0; s < S; ++s) {
                            uint32_t b = init_of(ring[k], s);
                            b = dmin(b, dadd(s == 0 ? lCur : cur[s - 1], 1.0f));
                            b = dmin(b, dadd(prev[s], 1.0f));
                            b = dmin(b, dadd(s == 0 ? prevL : prev[s - 1], kChamferDiag));
                            b = dmin(b, dadd(s == S - 1 ? prevR : prev[s + 1], kChamferDiag));
                            cur[s] = ((exists >> s) & 1u) ? b : kNone;
                        }
                    } else {
#pragma unroll
                        for (int s = S - 1; s >= 0; --s) {
                            uint32_t b = init_of(ring[k], s);
                            b = dmin(b, dadd(s == S - 1 ? rCur : cur[s + 1], 1.0f));
                            b = dmin(b, dadd(prev[s], 1.0f));
                            b = dmin(b, dadd(s == S - 1 ? prevR : prev[s + 1], kChamferDiag));
                            b = dmin(b, dadd(s == 0 ? prevL : prev[s - 1], kChamferDiag));
                            cur[s] = ((exists >> s) & 1u) ? b : kNone;
                        }
                    }
                    float* row = FWD ? pbase + r * dp : pbase - r * dp;
#pragma unroll
                    for (int g = 0; g < S / 4; ++g) {
                        f4v v;
                        v.x = __uint_as_float(cur[4 * g]); v.y = __uint_as_float(cur[4 * g + 1]);
                        v.z = __uint_as_float(cur[4 * g + 2]); v.w = __uint_as_float(cur[4 * g + 3]);
                        st_f4(row + 4 * g, v);
                    }
#pragma unroll
                    for (int s = 0; s < S; ++s) prev[s] = cur[s];
                    firstv[cb][t + 1] = cur[0];
                    lastv[cb][t + 1] = cur[S - 1];
                }
            }
            lds_barrier();  // neighbours exchange through LDS only
        }
        return true;
    };
    for (int j = 0;; j += 3) {
        if (!superstep(j, bufA, bufC)) break;
        if (!superstep(j + 1, bufB, bufA)) break;
        if (!superstep(j + 2, bufC, bufB)) break;
    }
}

// Threads per block for chunks of width w at strip width S.
inline int ChamferThreads(int w, int S) {
    const int cols = (w + S - 1) / S;
    return ((cols + 127) / 128) * 128;  // even and odd strips in separate wavefronts
}

// Strip width: `want` (4, 8 or 16; an autotuned choice) when a row of w
// columns fits one 1024-thread block at that width, else 8 columns (the
// measured best on 4096-wide chunks: fewer steps than 4 columns, shorter
// per-step chains than 16), 16 where 8 does not fit.
// I360HIP_CHAMFER_STRIP overrides both.
inline int ChamferStrip(int w, int want = 0) {
    int s = w <= 8192 ? 8 : 16;
    if ((want == 4 || want == 8 || want == 16) && w <= 1024 * want) s = want;
    if (const char* e = std::getenv("I360HIP_CHAMFER_STRIP")) {
        const int v = std::atoi(e);
        if ((v == 4 || v == 8 || v == 16) && w <= 1024 * v) s = v;
    }
    return s;
}

// log2(S) + 1 for the strip width a launch over chunks up to max_w wide
// picks: the shear of the distance planes it writes (ShearAt).
inline int ChamferShear(int max_w, int want_strip = 0) { return ShearShift(ChamferStrip(max_w, want_strip)); }

// Blocks per item of the shear pre-pass.
constexpr int kShearBlocks = 32;

// Shears the coverage boxes (unless the producer of the coverage already
// wrote the sheared copy, for this launch's ChamferShear, over the box),
// then launches the forward and backward passes over `n` work items
// (device table).  Every plane must hold ShearRows(w, h) rows of
// DistPitch(w) elements.
inline hipError_t launch_chamfer(const ChamferWork* d_work, int n, int max_w, hipStream_t s, int want_strip = 0,
                                 bool cov_sheared = false) {
    if (n <= 0) return hipSuccess;
    const int S = ChamferStrip(max_w, want_strip);
    const int nt = ChamferThreads(max_w, S);
    const dim3 sg(kShearBlocks, n);
    // rows per superstep: the backward pass buffers S distances per row in
    // three rotating buffers, so it takes fewer rows at wider strips
    if (S == 4) {
        if (!cov_sheared) chamfer_shear_kernel<4><<<sg, 256, 0, s>>>(d_work);
        chamfer_kernel<4, 8, true><<<n, nt, 0, s>>>(d_work);
        chamfer_kernel<4, 8, false><<<n, nt, 0, s>>>(d_work);
    } else if (S == 8) {
        if (!cov_sheared) chamfer_shear_kernel<8><<<sg, 256, 0, s>>>(d_work);
        chamfer_kernel<8, 4, true><<<n, nt, 0, s>>>(d_work);
        chamfer_kernel<8, 2, false><<<n, nt, 0, s>>>(d_work);
    } else {
        if (!cov_sheared) chamfer_shear_kernel<16><<<sg, 256, 0, s>>>(d_work);
        chamfer_kernel<16, 2, true><<<n, nt, 0, s>>>(d_work);
        chamfer_kernel<16, 1, false><<<n, nt, 0, s>>>(d_work);
    }
    return hipGetLastError();
}

}  // namespace i360
