// Coverage-distance (two-pass 3x3 chamfer) kernels shared by the RGB48 and
// float chunk renders.  Semantics: WindowsApp.Core/SourceFiles/
// SeamBlend.cpp:206-242, bit for bit.
//
// The serial raster loop makes pixel (x,y) of the forward pass depend on
// (x-1,y), (x,y-1), (x-1,y-1) and (x+1,y-1).  A block owns one
// (chunk, contributor) work item; thread t owns a strip of S columns and
// handles its row r at step r + t + 1, one workgroup barrier per step, so
// a pass takes rows + strips steps.  The left neighbours of a row were
// finished one step earlier.  The one operand that was not, the up-right
// neighbour of the strip's last column (the right strip's first column of
// the row above, which that strip computes in the same step), the strip
// recomputes itself from operands finished a step earlier: the right
// strip publishes min(init, up + 1, up-right + diag) of that pixel, and
// the strip adds its own two terms.  Every pixel, recomputed ones
// included, evaluates the min() of the same binary32 sums the serial loop
// does (min is exact and order-free), so the result equals the serial one
// bit for bit.  The backward pass mirrors the schedule (strips and columns
// in reverse order).
//
// Only the contributor's coverage box is processed: pixels outside it are
// uncovered, and an uncovered pixel's distance is 0 in both passes, so the
// box rows/columns are surrounded by known zeros (or by the chunk edge,
// where the serial loop has no neighbour either).  Distances outside the
// box are never written; readers treat them as 0.
//
// Each thread prefetches its next rows of input into registers, so the
// per-step critical path holds no global-memory latency.
//
// Sheared planes: at a given step, strip t works on row y0 + step - t - 1,
// so in a row-major plane the 64 lanes of a wavefront would touch 64
// different rows.  Distances (and a copy of the coverage the forward pass
// reads) are therefore stored sheared: pixel (x, y) lives at row
// y + x / S of its plane.  All strips of a wavefront then touch a single
// row at every step, in both passes, and their loads and stores coalesce.
#pragma once

#include <cstdlib>
#include <type_traits>

#include "common.hpp"

namespace i360 {

struct ChamferWork {
    const uint8_t* cov;  // w*h coverage (0/1), row pitch w; read only by chamfer_shear_kernel
    uint8_t* scov;       // sheared coverage over the box (chamfer_shear_kernel's copy, or the producer's)
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
// last column of the row pitch at the narrowest strip.
__host__ __device__ inline int ShearRows(int w, int h) { return h + (DistPitch(w) - 1) / ChamferMinStrip(w) + 1; }

// Element offset of pixel (x, y) in a sheared plane; shift = log2(S).
__host__ __device__ inline size_t ShearAt(int x, int y, int shift, int pitch) {
    return static_cast<size_t>(y + (x >> shift)) * pitch + x;
}
__host__ __device__ constexpr int ShearShift(int S) { return S == 4 ? 2 : (S == 8 ? 3 : 4); }

constexpr float kChamferDiag = 1.41421356f;  // SeamBlend.cpp:209

typedef float f4v __attribute__((ext_vector_type(4)));
typedef uint32_t u2v __attribute__((ext_vector_type(2)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t ld_u8(const uint8_t* p) { return *gptr(p); }
__device__ __forceinline__ f4v ld_f4(const float* p) { return *gptr(reinterpret_cast<const f4v*>(p)); }
__device__ __forceinline__ void st_f4(float* p, f4v v) { *gptr(reinterpret_cast<f4v*>(p)) = v; }

// Copies each item's coverage box, widened to whole 16-column groups, into
// its sheared plane, G = min(S, 8) bytes per thread (a G-column group at a
// multiple of G never straddles two shear rows); columns at or past w are
// written as 0.  Grid: (blocks per item, items).
template <int S>
__global__ void __launch_bounds__(256) chamfer_shear_kernel(const ChamferWork* __restrict__ work) {
    constexpr int G = S < 8 ? S : 8;
    const ChamferWork wk = work[blockIdx.y];
    const int x0 = wk.box[0], y0 = wk.box[1], x1 = wk.box[2], y1 = wk.box[3];
    if (x1 < 0) return;
    const int w = wk.w, cp = DistPitch(w);
    const int ga = x0 & ~15, gb = min((x1 + 16) & ~15, cp);
    const int ng = (gb - ga) / G;
    const size_t total = static_cast<size_t>(ng) * (y1 - y0 + 1);
    const bool aligned = (w & (G - 1)) == 0;
    for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int y = y0 + static_cast<int>(e / ng), x = ga + G * static_cast<int>(e % ng);
        const uint8_t* c = wk.cov + static_cast<size_t>(y) * w + x;
        uint8_t* d = wk.scov + ShearAt(x, y, ShearShift(S), cp);
        uint32_t v[G / 4] = {};
        if (aligned && x + G <= w) {
#pragma unroll
            for (int g = 0; g < G / 4; ++g) v[g] = *gptr(reinterpret_cast<const uint32_t*>(c) + g);
        } else {
#pragma unroll
            for (int k = 0; k < G; ++k)
                if (x + k < w) v[k >> 2] |= ld_u8(c + k) << (8 * (k & 3));
        }
        if constexpr (G == 8) {
            u2v q;
            q.x = v[0]; q.y = v[1];
            *gptr(reinterpret_cast<u2v*>(d)) = q;
        } else {
            *gptr(reinterpret_cast<uint32_t*>(d)) = v[0];
        }
    }
}

// Block layout: thread t owns strip t (64 consecutive strips per
// wavefront).
//
// Neighbours beyond the processed area are constants: a pixel that exists
// outside the coverage box is uncovered (0), a pixel beyond the chunk edge
// does not exist (+inf, which no min picks).  They come from a sentinel
// LDS slot (before the first strip in processing order), per-thread
// constants for the row before the box, and a per-thread constant for the
// column past the last strip.  Every LDS read is therefore unconditional
// and a pixel's update is branch-free.  Distances are non-negative floats,
// whose order is their bit patterns' unsigned order: min is one integer
// instruction.
__device__ __forceinline__ uint32_t dmin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__host__ __device__ constexpr int RWmax(int S) { return S / 4; }
__device__ __forceinline__ uint32_t dadd(uint32_t a, float c) { return __float_as_uint(__uint_as_float(a) + c); }

// Terms of the passes, in processing order: FWD runs left to right, top
// to bottom; BWD right to left, bottom to top.  "Upstream" is the
// neighbour strip a row's values flow in from (left in FWD, right in BWD),
// "downstream" the other one.  Column i of the processing order is strip
// column FWD ? i : S - 1 - i.
template <int S, int P, bool FWD>
__global__ void __launch_bounds__(1024) chamfer_kernel(const ChamferWork* __restrict__ work) {
    // Written at step T into buffer T & 1, read at step T + 1; slot t + 1
    // holds strip t, slot 0 / nact + 1 the upstream sentinel (FWD / BWD).
    //   edgev: the strip's last column (processing order) of its latest row
    //   partv: for the strip's first column of its next row, the min of that
    //          pixel's initial value and its two terms from the latest row
    __shared__ uint32_t edgev[2][1024 + 2];
    __shared__ uint32_t partv[2][1024 + 2];
    const ChamferWork wk = work[blockIdx.x];
    const int x0 = wk.box[0], y0 = wk.box[1], x1 = wk.box[2], y1 = wk.box[3];
    if (x1 < 0) return;  // block-uniform
    const int w = wk.w, h = wk.h;
    const int xa = (x0 / S) * S;
    const int nact = (x1 + 1 - xa + S - 1) / S;
    const int rows = y1 - y0 + 1;
    const int t = threadIdx.x;  // strip index
    const bool active = t < nact;
    const int xs = xa + t * S;
    const int dp = DistPitch(w);
    const int xl = xa + min(t, nact - 1) * S;  // load strip: inactive threads load the last active one
    constexpr int SH = ShearShift(S);
    constexpr uint32_t kNone = 0x7f800000u;  // +inf: no pixel there
    const float kInf = static_cast<float>(w + h);
    const int lag = FWD ? t : (nact - 1 - t);
    const bool edgePrev = FWD ? (y0 > 0) : (y1 < h - 1);  // a row exists before the box (all uncovered)
    const int upSlot = FWD ? t : t + 2, dnSlot = FWD ? t + 2 : t;
    const bool lastStrip = FWD ? (t == nact - 1) : (t == 0);  // no downstream strip
    auto col = [](int i) { return FWD ? i : S - 1 - i; };

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
    // the row before the box: the columns left and right of the strip there
    const uint32_t pL0 = (edgePrev && xs > 0) ? 0u : kNone;
    const uint32_t pR0 = (edgePrev && xs + S < w) ? 0u : kNone;
    const uint32_t pUp0 = FWD ? pL0 : pR0, pDn0 = FWD ? pR0 : pL0;
    // the columns left and right of the processed strips, in the box rows
    const uint32_t sl = xa > 0 ? 0u : kNone, sr = (xa + nact * S < w) ? 0u : kNone;
    const uint32_t sDn = FWD ? sr : sl;
    if (threadIdx.x == 0) {
        const int slot = FWD ? 0 : nact + 1;
        edgev[0][slot] = FWD ? sl : sr;
        edgev[1][slot] = FWD ? sl : sr;
    }
    // Steady state: a wavefront whose strips are all full interior strips
    // (active, every column inside the box and the chunk, a downstream
    // strip present) runs the supersteps in which each of its strips
    // computes a row 1 .. rows - 2 without any of the edge cases.
    constexpr uint32_t kAll = (1u << S) - 1u;
    const bool interior = active && !lastStrip && inBox == kAll && exists == kAll;
    const bool waveFast = __all(interior);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lagMin = FWD ? 64 * wv : nact - 64 - 64 * wv, lagMax = lagMin + 63;

    // Sheared plane rows: at step T every strip loads its row T - lag and
    // stores its row T - lag - 1, and under the shear these are one plane
    // row for the whole block, so addresses are a block-uniform row plus
    // the strip's column.  A load row is clamped into the plane: where it
    // falls outside, no strip's row is inside the box and the data is not
    // used (inactive threads load the last active strip's columns).
    const int srows = ShearRows(w, h), xaS = xa / S;
    auto loadRow = [&](int T) {
        const int q = FWD ? (y0 + T + xaS) : (y1 - T + nact - 1 + xaS);
        return min(max(q, 0), srows - 1);
    };
    auto storeRow = [&](int T) { return FWD ? (y0 + T - 1 + xaS) : (y1 - T + nact + xaS); };

    // raw prefetched row data; converted only at the row's step so that no
    // instruction depends on a load until then
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
    // distances of the strip's latest row (initially the row before the box)
    uint32_t prev[S];
#pragma unroll
    for (int s = 0; s < S; ++s) prev[s] = (edgePrev && ((exists >> s) & 1u)) ? 0u : kNone;
    uint32_t initCur[S];           // initial values of the row being computed
#pragma unroll
    for (int s = 0; s < S; ++s) initCur[s] = kNone;
    uint32_t prevEdge = pUp0;      // upstream strip's last column, row before the latest
    uint32_t pprevLast = kNone;    // this strip's last column, row before the latest

    auto load_row = [&](int T, uint32_t (&dst)[RW]) {
        const size_t ro = static_cast<size_t>(loadRow(T)) * dp;
        if constexpr (FWD) {
            const uint8_t* c = wk.scov + ro + xl;
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
            const float* d = wk.dist + ro + xl;
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
    // distance (BWD); outside the box 0.  The interior form skips the box
    // mask.
    auto init_of = [&](const uint32_t (&src)[RW], int s, auto interiorc) -> uint32_t {
        constexpr bool F = decltype(interiorc)::value;
        if constexpr (FWD) {
            // coverage byte (0 or 1; 0 outside the box) times w + h
            const uint32_t word = F ? src[s >> 2] : (src[s >> 2] & byteMask[s >> 2]);
            const float c = static_cast<float>((word >> (8 * (s & 3))) & 0xffu);
            return __float_as_uint(c * kInf);
        } else {
            if constexpr (F) return src[s];
            return ((inBox >> s) & 1u) ? src[s] : 0u;
        }
    };
    // step T: row T - lag - 1 is computed, and row T - lag (held in slot
    // T mod P of superstep T / P's buffer) supplies the initial values of
    // the next row
    auto fetch = [&](int j, uint32_t (&dst)[P][RW]) {
#pragma unroll
        for (int k = 0; k < P; ++k) load_row(P * j + k, dst[k]);
    };
    fetch(0, bufA);
    fetch(1, bufB);
    lds_barrier();  // sentinels visible

    // one step; the interior form assumes what the steady-state test
    // guarantees (every strip active and interior, 1 <= r <= rows - 2)
    auto step = [&](auto interiorc, int T, const uint32_t (&raw)[RW]) {
        constexpr bool F = decltype(interiorc)::value;
        const int r = T - lag - 1;
        if (F || (active && r >= -1 && r < rows)) {
            uint32_t initNext[S];
#pragma unroll
            for (int s = 0; s < S; ++s) initNext[s] = init_of(raw, s, interiorc);
            if (F || r >= 0) {
                const int rb = (T - 1) & 1;
                // neighbours, read together
                const uint32_t eUp = edgev[rb][upSlot];
                const uint32_t pDn = partv[rb][dnSlot];
                const uint32_t diagUp = (F || r != 0) ? prevEdge : pUp0;
                // downstream strip's first column, row above: recomputed
                // from the terms it published and this strip's two
                uint32_t ur = dmin(dmin(pDn, dadd(prev[col(S - 1)], 1.0f)), dadd(pprevLast, kChamferDiag));
                if constexpr (!F) ur = r == 0 ? pDn0 : (lastStrip ? sDn : ur);
                uint32_t cur[S];
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    const int s = col(i);
                    uint32_t b = initCur[s];
                    b = dmin(b, dadd(prev[s], 1.0f));
                    b = dmin(b, dadd(i == 0 ? diagUp : prev[col(i - 1)], kChamferDiag));
                    b = dmin(b, dadd(i == S - 1 ? ur : prev[col(i + 1)], kChamferDiag));
                    b = dmin(b, dadd(i == 0 ? eUp : cur[col(i - 1)], 1.0f));
                    cur[s] = (F || ((exists >> s) & 1u)) ? b : kNone;
                }
                // BWD rewrites the plane in place: where the strip's four
                // values equal what the forward pass left (initCur: the
                // forward bits, or 0 outside the box, where the forward
                // pass stored 0), the plane already holds the result
                bool same = !FWD;
                if constexpr (!FWD) {
#pragma unroll
                    for (int s = 0; s < S; ++s) same = same && cur[s] == initCur[s];
                }
                if (!same) {
                    float* row = wk.dist + static_cast<size_t>(storeRow(T)) * dp + xs;
#pragma unroll
                    for (int g = 0; g < S / 4; ++g) {
                        f4v v;
                        v.x = __uint_as_float(cur[4 * g]); v.y = __uint_as_float(cur[4 * g + 1]);
                        v.z = __uint_as_float(cur[4 * g + 2]); v.w = __uint_as_float(cur[4 * g + 3]);
                        st_f4(row + 4 * g, v);
                    }
                }
                pprevLast = prev[col(S - 1)];
#pragma unroll
                for (int s = 0; s < S; ++s) prev[s] = cur[s];
                prevEdge = eUp;
            }
            // publish (row -1: the constants of the row before the box)
            edgev[T & 1][t + 1] = prev[col(S - 1)];
            partv[T & 1][t + 1] =
                dmin(dmin(initNext[col(0)], dadd(prev[col(0)], 1.0f)), dadd(prev[col(1)], kChamferDiag));
#pragma unroll
            for (int s = 0; s < S; ++s) initCur[s] = initNext[s];
        }
    };

    const int tmax = rows + nact - 1;
    // one superstep; false (block-uniform) once past the last step
    auto superstep = [&](int j, const uint32_t (&ring)[P][RW], uint32_t (&ld)[P][RW]) -> bool {
        const int T0 = P * j;
        if (T0 > tmax) return false;
        fetch(j + 2, ld);
        if (waveFast && T0 - lagMax - 1 >= 1 && T0 + P - 1 - lagMin - 1 <= rows - 2) {
            // wave-uniform; all P steps exist (T0 + P - 1 <= rows - 1 + lagMin <= tmax)
#pragma unroll
            for (int k = 0; k < P; ++k) {
                step(std::true_type{}, T0 + k, ring[k]);
                lds_barrier();
            }
        } else {
            const int nsteps = min(P, tmax - T0 + 1);  // uniform
#pragma unroll
            for (int k = 0; k < P; ++k) {
                if (k >= nsteps) continue;  // uniform: all threads skip the barrier together
                step(std::false_type{}, T0 + k, ring[k]);
                lds_barrier();  // neighbours exchange through LDS only
            }
        }
        return true;
    };
    for (int j = 0;; j += 3) {
        if (!superstep(j, bufA, bufC)) break;
        if (!superstep(j + 1, bufB, bufA)) break;
        if (!superstep(j + 2, bufC, bufB)) break;
    }
}

// Threads per block for chunks of width w at strip width S (whole
// wavefronts).
inline int ChamferThreads(int w, int S) {
    const int cols = (w + S - 1) / S;
    return ((cols + 63) / 64) * 64;
}

// Strip width: `want` (4, 8 or 16; an autotuned choice) when a row of w
// columns fits one 1024-thread block at that width, else the narrowest
// strip that fits (4 columns measured best for the four-group gigapixel
// step on 4096-wide chunks: 8 columns take about as long on their own, but
// their longer per-step chains overlap the other groups' stages worse).
// I360HIP_CHAMFER_STRIP overrides both.
inline int ChamferStrip(int w, int want = 0) {
    int s = ChamferMinStrip(w);
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
        chamfer_kernel<4, 4, false><<<n, nt, 0, s>>>(d_work);
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
