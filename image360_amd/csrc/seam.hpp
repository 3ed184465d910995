// Pairwise DP seam refinement (WindowsApp.Core/SourceFiles/
// SeamBlend.cpp:293-445), shared by the RGB48 and float renders.
//
// Three launches per pair (or per round of pairs in the batched form):
//  scan     one wavefront per line: the crossover of the current owner map
//           and the contested band around it (ballots), then the band's
//           pixel costs |I - J|, written to a per-line cost row;
//  dp       one block per pair: the line-to-line DP over band positions
//           lo+k, four lanes per position, each lane a quarter of the
//           previous line's candidates (at most 49, read from LDS, a tree
//           minimum), the quarters' minima joined across lanes with a
//           strict <, so the first minimum wins as in the reference's
//           serial scan; then the backtrack through back-pointer rows held
//           in registers;
//  reassign one wavefront per line: I/J-owned pixels inside each contested
//           band go to the side of the chosen seam.
// Img supplies rgb(p, M[3]) and the smoothness weight per pixel of seam
// movement (2 in RGB48 units, SeamBlend.cpp:309; 2/65535 for [0,1] images).
#pragma once

#include "chamfer.hpp"
#include "common.hpp"

namespace i360 {

template <typename M> struct SeamMax;
template <> struct SeamMax<float> { static constexpr float value = 3.402823466e+38f; };
template <> struct SeamMax<double> { static constexpr double value = 1.7976931348623157e+308; };

__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_value(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// RGB48 contributor (reference storage).
struct U16Img {
    const uint16_t* p;
    __device__ __forceinline__ void rgb(size_t i, float (&o)[3]) const {
        const auto* g = gptr(p);
        o[0] = static_cast<float>(g[i * 3]);
        o[1] = static_cast<float>(g[i * 3 + 1]);
        o[2] = static_cast<float>(g[i * 3 + 2]);
    }
    __device__ static constexpr float smooth() { return 2.0f; }
};

constexpr int kSeamRadius = 24;  // SeamBlend.cpp:319
constexpr int kDpWaves = 4;      // wavefronts of the DP kernel (one per SIMD)
constexpr int kDpThreads = 64 * kDpWaves;
constexpr int kDpLines = 16;     // cost rows staged in LDS per batch

__device__ __forceinline__ size_t seam_at(bool rowScan, int ox0, int oy0, int w, int line, int t) {
    return rowScan ? static_cast<size_t>(oy0 + line) * w + (ox0 + t) : static_cast<size_t>(oy0 + t) * w + (ox0 + line);
}


// Phase 1 of a pair: for every line, the crossover of the current owner map
// between the pair and the contested band around it (SeamBlend.cpp:325-360).
// One wavefront per line, lanes along the line, four 64-pixel segments in
// flight per lane; blocks cover lines, so all lines of a pair (and all
// pairs of a round) scan in parallel.
// The band's pixel costs sum_c |I_c - J_c| (SeamBlend.cpp:371-378) go to
// costs[line * 64 + k] for band position lo + k.
// covI / covJ: the contributors' coverage bytes, row-major (covShift < 0)
// or in the chamfer's sheared planes (covShift = the launch's shear).
template <class Img, typename M, typename Own, int U = 4>
__device__ __forceinline__ void seam_scan_body(Img imgI, Img imgJ, const uint8_t* __restrict__ covI,
                                               const uint8_t* __restrict__ covJ, int covShift,
                                               const int* __restrict__ boxes,
                                               int ci, int cj, int w, const Own* __restrict__ owner,
                                               int* __restrict__ lohi, int* __restrict__ chosen,
                                               int* __restrict__ contested, M* __restrict__ costs, int line) {
    const int* bI = boxes + 4 * ci;
    const int* bJ = boxes + 4 * cj;
    const auto* gbI = gptr(bI);
    const auto* gbJ = gptr(bJ);
    if (gbI[2] < 0 || gbJ[2] < 0) return;
    const int ox0 = max(gbI[0], gbJ[0]), oy0 = max(gbI[1], gbJ[1]);
    const int ox1 = min(gbI[2], gbJ[2]), oy1 = min(gbI[3], gbJ[3]);
    if (ox0 > ox1 || oy0 > oy1) return;
    const int bw = ox1 - ox0 + 1, bh = oy1 - oy0 + 1;
    const bool rowScan = bw >= bh;
    const int lines = rowScan ? bh : bw, len = rowScan ? bw : bh;
    if (line >= lines) return;  // wave-uniform
    const int lane = threadIdx.x & 63;
    const auto* gown = gptr(owner);
    int first = -1, cross = -1;
    for (int b0 = 0; b0 < len && cross < 0; b0 += 64 * U) {
        int ob[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = b0 + 64 * u + lane;
            ob[u] = (t < len) ? gown[seam_at(rowScan, ox0, oy0, w, line, t)] : -2;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (cross >= 0) break;
            const int base = b0 + 64 * u;
            const int o = ob[u];
            const bool ij = (o == ci || o == cj);
            const unsigned long long m = __ballot(ij);
            int start = 0;
            if (first < 0) {
                if (m == 0ull) continue;
                const int f = __ffsll(static_cast<long long>(m)) - 1;
                first = __shfl(o, f);
                start = f + 1;
            }
            unsigned long long m2 = __ballot(ij && o != first);
            m2 = (start >= 64) ? 0ull : (m2 & (~0ull << start));
            if (m2) cross = base + __ffsll(static_cast<long long>(m2)) - 1;
        }
    }
    int lo = 1, hi = 0;
    if (cross >= 0) {
        const int a = max(0, cross - kSeamRadius), b = min(len - 1, cross + kSeamRadius);
        const int t = a + lane;
        bool both = false;
        if (t <= b) {
            const size_t p = covShift < 0 ? seam_at(rowScan, ox0, oy0, w, line, t)
                                          : ShearAt(rowScan ? ox0 + t : ox0 + line, rowScan ? oy0 + line : oy0 + t,
                                                    covShift, DistPitch(w));
            both = gptr(covI)[p] && gptr(covJ)[p];
        }
        const unsigned long long m = __ballot(both);
        if (m) {
            lo = a + __ffsll(static_cast<long long>(m)) - 1;
            hi = a + 63 - __clzll(static_cast<long long>(m));
        }
    }
    if (lane <= hi - lo) {
        const size_t p = seam_at(rowScan, ox0, oy0, w, line, lo + lane);
        M x[3], y[3];
        imgI.rgb(p, x);
        imgJ.rgb(p, y);
        M cost = M(0);
#pragma unroll
        for (int c = 0; c < 3; ++c) cost = cost + fabs(x[c] - y[c]);
        gptr(costs)[static_cast<size_t>(line) * 64 + lane] = cost;
    }
    if (lane == 0) {
        gptr(lohi)[2 * line] = lo;
        gptr(lohi)[2 * line + 1] = hi;
        gptr(chosen)[line] = -1;
        if (hi >= lo) atomicAdd(contested, 1);
    }
}

// Lines per block of the scan kernel (one wavefront each).
constexpr int kScanLinesPerBlock = 4;

template <class Img, typename M>
__global__ void __launch_bounds__(256) seam_scan_kernel(Img imgI, Img imgJ, const uint8_t* __restrict__ covI,
                                                        const uint8_t* __restrict__ covJ,
                                                        const int* __restrict__ boxes, int ci, int cj, int w,
                                                        const int32_t* __restrict__ owner, int* __restrict__ lohi,
                                                        int* __restrict__ chosen, int* __restrict__ contested,
                                                        M* __restrict__ costs) {
    const int line = blockIdx.x * kScanLinesPerBlock + (threadIdx.x >> 6);
    seam_scan_body<Img, M, int32_t>(imgI, imgJ, covI, covJ, -1, boxes, ci, cj, w, owner, lohi, chosen, contested, costs,
                                    line);
}

// Overlap of the pair's coverage boxes and its scan direction (the
// reference's rowScan, SeamBlend.cpp:313-323); false when the boxes do
// not overlap.
struct SeamGeom {
    int ox0, oy0, lines;
    bool rowScan, iLow;
};
__device__ __forceinline__ bool seam_geom(const int* __restrict__ boxes, int ci, int cj, SeamGeom& g) {
    const auto* bI = gptr(boxes + 4 * ci);
    const auto* bJ = gptr(boxes + 4 * cj);
    if (bI[2] < 0 || bJ[2] < 0) return false;  // an empty coverage box
    g.ox0 = max(bI[0], bJ[0]);
    g.oy0 = max(bI[1], bJ[1]);
    const int ox1 = min(bI[2], bJ[2]), oy1 = min(bI[3], bJ[3]);
    if (g.ox0 > ox1 || g.oy0 > oy1) return false;
    const int bw = ox1 - g.ox0 + 1, bh = oy1 - g.oy0 + 1;
    g.rowScan = bw >= bh;
    const float cI = g.rowScan ? static_cast<float>(bI[0] + bI[2]) * 0.5f : static_cast<float>(bI[1] + bI[3]) * 0.5f;
    const float cJ = g.rowScan ? static_cast<float>(bJ[0] + bJ[2]) * 0.5f : static_cast<float>(bJ[1] + bJ[3]) * 0.5f;
    g.iLow = cI <= cJ;
    g.lines = g.rowScan ? bh : bw;
    return true;
}

// LDS of the DP kernel.
template <typename M>
struct SeamDpShared {
    static constexpr int kSpan = (2 * kSeamRadius + 1 + 3) / 4;  // candidates per lane: 4 groups cover a band
    M cost[2][kDpLines][64];                   // cost rows of a batch (double buffered)
    M dpv[2][64];                              // dp row of the last two contested lines (by position)
    __align__(16) uint8_t bk[2][kDpLines][64];  // back-pointer rows of a batch
    int prv[2][kDpLines];                      // predecessor line of each line of a batch
    M dpfinal[64];                             // dp row of the last contested line
};

// Lanes l and l ^ S (S = 16: adjacent rows, 32: halves of the wavefront)
// exchange a dword: lo gets the lower lane's value, hi the upper lane's, in
// both lanes (gfx950 permlane swaps, no LDS).
template <int S>
__device__ __forceinline__ void lane_pair(int x, int& lo, int& hi) {
    if constexpr (S == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        lo = static_cast<int>(r[0]);
        hi = static_cast<int>(r[1]);
    } else {
        static_assert(S == 32, "rows or halves");
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        lo = static_cast<int>(r[0]);
        hi = static_cast<int>(r[1]);
    }
}
template <int S>
__device__ __forceinline__ void lane_pair(float x, float& lo, float& hi) {
    int a, b;
    lane_pair<S>(__float_as_int(x), a, b);
    lo = __int_as_float(a);
    hi = __int_as_float(b);
}
template <int S>
__device__ __forceinline__ void lane_pair(double x, double& lo, double& hi) {
    const long long v = __double_as_longlong(x);
    int l0, h0, l1, h1;
    lane_pair<S>(static_cast<int>(v & 0xffffffffll), l0, h0);
    lane_pair<S>(static_cast<int>(v >> 32), l1, h1);
    lo = __longlong_as_double((static_cast<long long>(l1) << 32) | static_cast<unsigned int>(l0));
    hi = __longlong_as_double((static_cast<long long>(h1) << 32) | static_cast<unsigned int>(h0));
}
// The minimum of lanes l and l ^ S, in both; the lower lane holds the lower
// candidate indices and wins ties (the upper one only on a strict <).
template <int S, typename M>
__device__ __forceinline__ void lane_pair_min(M& v, int& k) {
    M vl, vh;
    int kl, kh;
    lane_pair<S>(v, vl, vh);
    lane_pair<S>(k, kl, kh);
    const bool up = vh < vl;
    v = up ? vh : vl;
    k = up ? kh : kl;
}

// Phase 2 for wavefront W: the line-to-line DP (SeamBlend.cpp:362-399).
//
// A band is at most 2 * kSeamRadius + 1 wide.  Wavefront W holds target
// positions t = 16 W + (lane & 15), four lanes each: row g = lane >> 4 of
// the wavefront evaluates the previous line's candidates pk = kSpan g + i,
// read from LDS, with no branch: a candidate outside the last line's band
// holds +inf, which never wins a strict <.  The penalty
// smooth * |t - (lastLo + pk)| is formed from (t - lastLo - kSpan g) - i,
// exact in M for these small integers, so each candidate's value is the
// reference's bit for bit.  A lane's kSpan values are reduced as a tree
// whose right operand (higher candidates) wins only on a strict <, and the
// four rows' minima likewise across lane pairs (rows, then halves), so the
// first minimum wins as in the reference's serial scan.  Row 0 writes the
// new dp value and back-pointer of its position; one barrier per line.
//
// Cost rows of kDpLines lines are staged in LDS and their bands held in
// registers (lane 2l / 2l+1: lo / hi of line l), double buffered: the next
// batch is loaded while the current one runs.  Back-pointer rows are
// collected in LDS and written out at the batch's end.
template <class Img, typename M, int W>
__device__ __forceinline__ void seam_dp_forward(SeamDpShared<M>& sh, int lines, const M* __restrict__ costs,
                                                const int* __restrict__ lohi, int* __restrict__ prevl,
                                                uint8_t* __restrict__ back, int& last_out, int& lastW_out) {
    constexpr int kSpan = SeamDpShared<M>::kSpan;
    static_assert(kDpWaves == 4, "16 target positions per wavefront");
    static_assert(4 * kSpan >= 2 * kSeamRadius + 1 && 4 * kSpan <= 64, "candidate split");
    constexpr int kPer = kDpLines * 64 / kDpThreads;
    static_assert(2 * kDpLines <= 64, "bands of a batch fit one wavefront");
    static_assert(kDpLines * 16 <= kDpThreads, "one dword of back-pointers per thread");
    const int lane = threadIdx.x & 63;
    const int t = 16 * W + (lane & 15);  // target position of this lane
    const int g0 = kSpan * (lane >> 4);  // first candidate of this lane's row
    const bool row0 = lane < 16;         // writes position t
    const M kOut = __builtin_huge_val();  // dp value of a position outside the band
    M creg[kPer];
    int bnext = 0;
    auto fetch = [&](int b0) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = static_cast<int>(threadIdx.x) + u * kDpThreads;
            const int l = min(b0 + (idx >> 6), lines - 1);
            creg[u] = gptr(costs)[static_cast<size_t>(l) * 64 + (idx & 63)];
        }
        if (lane < 2 * kDpLines) bnext = gptr(lohi)[2 * min(b0 + (lane >> 1), lines - 1) + (lane & 1)];
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = static_cast<int>(threadIdx.x) + u * kDpThreads;
            sh.cost[buf][idx >> 6][idx & 63] = creg[u];
        }
    };
    // rows of uncontested lines carry stale bytes: the backtrack never
    // visits them (it follows the predecessor chain of contested lines)
    auto flush = [&](int buf, int b0, int nb) {
        const int tid = static_cast<int>(threadIdx.x);
        if (tid < nb * 16)
            reinterpret_cast<I360_GLOBAL uint32_t*>(gptr(back) + static_cast<size_t>(b0) * 64)[tid] =
                reinterpret_cast<const uint32_t*>(&sh.bk[buf][0][0])[tid];
        if (tid < nb) gptr(prevl)[b0 + tid] = sh.prv[buf][tid];
    };
    fetch(0);
    commit(0);
    int bcur = bnext;
    __syncthreads();
    int last = -1, lastLo = 0, lastW = 0, ord = 0;
    M dpv = kOut;  // dp value of the last contested line at position t
    // penalties smooth * |(lo + t) - (lastLo + g0 + i)| of this lane's
    // candidates, which depend on the line only through the band shift
    // lo - lastLo
    M pen[kSpan] = {};
    int pshift = INT_MIN;
    for (int b0 = 0, bi = 0; b0 < lines; b0 += kDpLines, ++bi) {
        const int buf = bi & 1;
        const bool more = b0 + kDpLines < lines;
        if (more) fetch(b0 + kDpLines);
        const int nb = min(kDpLines, lines - b0);
        for (int l = 0; l < nb;) {
            // Lines run in stretches of one band shift lo - lastLo: the
            // penalties are formed at the head of a stretch and stay fixed
            // (loop-invariant, so they are not copied line to line) until a
            // contested line's shift differs.
            const int lo0 = __builtin_amdgcn_readlane(bcur, 2 * l);
            if (last >= 0 && lo0 - lastLo != pshift) {  // uniform
                pshift = lo0 - lastLo;
                const M d0 = static_cast<M>(pshift + t - g0);
#pragma unroll
                for (int i = 0; i < kSpan; ++i) pen[i] = Img::smooth() * fabs(d0 - static_cast<M>(i));
            }
            for (; l < nb; ++l) {
                const int lo = __builtin_amdgcn_readlane(bcur, 2 * l);
                const int hi = __builtin_amdgcn_readlane(bcur, 2 * l + 1);
                if (hi < lo) continue;  // block-uniform
                if (last >= 0 && lo - lastLo != pshift) break;  // uniform: a new stretch
                const int lw = hi - lo + 1;
                const M c = sh.cost[buf][l][t];
                const int pb = ord & 1;
                ++ord;
                M v = kOut;
                int bp = -1;
                if (last < 0) {
                    v = (t < lw) ? c : kOut;
                } else {
                    const M* prev = &sh.dpv[pb ^ 1][g0];
                    M cv[kSpan];
                    int ck[kSpan];
#pragma unroll
                    for (int i = 0; i < kSpan; ++i) {
                        cv[i] = prev[i] + pen[i];
                        ck[i] = i;
                    }
#pragma unroll
                    for (int st = 1; st < kSpan; st *= 2)
#pragma unroll
                        for (int i = 0; i + st < kSpan; i += 2 * st) {
                            ck[i] = (cv[i + st] < cv[i]) ? ck[i + st] : ck[i];
                            cv[i] = __builtin_fmin(cv[i], cv[i + st]);  // no NaN; a tie keeps the same bits
                        }
                    M best = cv[0];
                    bp = g0 + ck[0];
                    lane_pair_min<16>(best, bp);
                    lane_pair_min<32>(best, bp);
                    v = (t < lw) ? c + best : kOut;
                }
                if (row0) {
                    sh.dpv[pb][t] = v;
                    sh.bk[buf][l][t] = static_cast<uint8_t>(bp & 0xFF);
                    dpv = v;
                }
                if (W == 0 && lane == 0) sh.prv[buf][l] = last;
                lds_barrier();  // the new dp row is visible to every wavefront
                last = b0 + l;
                lastLo = lo;
                lastW = lw;
            }
        }
        if (more) commit(buf ^ 1);
        bcur = bnext;
        __syncthreads();
        flush(buf, b0, nb);
    }
    if (row0) sh.dpfinal[t] = dpv;
    last_out = last;
    lastW_out = lastW;
}

// DP and backtrack of one pair; blockDim.x == kDpThreads.  Writes the
// chosen seam position of every contested line (SeamBlend.cpp:362-407).
template <class Img, typename M>
__device__ __forceinline__ void seam_dp_body(const int* __restrict__ contested, const int* __restrict__ boxes,
                                             int ci, int cj, const M* __restrict__ costs,
                                             const int* __restrict__ lohi, int* __restrict__ prevl,
                                             int* __restrict__ chosen, uint8_t* __restrict__ back) {
    SeamGeom g;
    if (!seam_geom(boxes, ci, cj, g)) return;
    const int lines = g.lines;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    // none contested: no DP and no reassignment (SeamBlend.cpp:409)
    if (*gptr(contested) == 0) return;
    __shared__ SeamDpShared<M> sh;
    static_assert(kDpWaves == 4, "one forward instance per wavefront");
    int last = -1, lastW = 0;
    // each wavefront runs its own instance (candidate lanes as constants);
    // all execute the same barriers in the same order
    switch (wave) {
        case 0: seam_dp_forward<Img, M, 0>(sh, lines, costs, lohi, prevl, back, last, lastW); break;
        case 1: seam_dp_forward<Img, M, 1>(sh, lines, costs, lohi, prevl, back, last, lastW); break;
        case 2: seam_dp_forward<Img, M, 2>(sh, lines, costs, lohi, prevl, back, last, lastW); break;
        default: seam_dp_forward<Img, M, 3>(sh, lines, costs, lohi, prevl, back, last, lastW); break;
    }
    if (last < 0) return;  // no contested pixel in this overlap (block-uniform)
    __syncthreads();  // the last line's dp row in LDS, back-pointer rows and predecessors in memory
    if (wave != 0) return;  // the backtrack is one wavefront's: no barrier follows

    // Phase 3: backtrack from the best end position.  The wavefront holds the
    // back-pointer bytes of 64 lines in registers (register j: line b0 + j,
    // lane: position) and each line's predecessor and band start in lane j,
    // and follows the chain with readlane, so a step costs no memory round
    // trip; the next 64 lines load while a batch is walked.
    const int lane = static_cast<int>(threadIdx.x);
    int k = 0;
    {
        const M f = lane < lastW ? sh.dpfinal[lane] : M(0);
        M best = SeamMax<M>::value;
        for (int q = 0; q < lastW; ++q) {  // the first minimum (strict <)
            const M v = lane_value(f, q);
            if (v < best) {
                best = v;
                k = q;
            }
        }
    }
    int bkr[64], nbkr[64];
    int prv = 0, lo0 = 0, nprv = 0, nlo0 = 0;
    auto load = [&](int b0, int (&r)[64], int& pv, int& lv) {  // rows b0.. (<= last)
#pragma unroll
        for (int j = 0; j < 64; ++j) r[j] = gptr(back)[static_cast<size_t>(min(b0 + j, last)) * 64 + lane];
        const int l = min(b0 + lane, last);
        pv = gptr(prevl)[l];
        lv = gptr(lohi)[2 * l];
    };
    int line = last;
    int b0 = line & ~63;
    load(b0, nbkr, nprv, nlo0);
    while (true) {  // uniform; line >= 0 and in [b0, b0 + 64), whose rows are in nbkr
        // the walked batch is copied out of the load targets first, so no
        // load is in flight for it while the next batch loads
#pragma unroll
        for (int j = 0; j < 64; ++j) bkr[j] = nbkr[j];
        prv = nprv;
        lo0 = nlo0;
        const int nb0 = b0 - 64;
        if (nb0 >= 0) load(nb0, nbkr, nprv, nlo0);
        int ch = -1;
#pragma unroll
        for (int j = 63; j >= 0; --j) {
            if (line == b0 + j) {  // uniform: line b0 + j is on the chain
                ch = lane == j ? __builtin_amdgcn_readlane(lo0, j) + k : ch;
                k = __builtin_amdgcn_readlane(bkr[j], k & 63);
                line = __builtin_amdgcn_readlane(prv, j);
            }
        }
        if (ch >= 0) gptr(chosen)[b0 + lane] = ch;
        if (line < 0) break;
        b0 = line & ~63;
        if (b0 != nb0) load(b0, nbkr, nprv, nlo0);  // the chain skipped a batch of uncontested lines
    }
}

// Phase 4 for one line (one wavefront): reassign I/J-owned pixels inside
// the contested band to the side of the chosen seam (SeamBlend.cpp:409-437).
template <typename Own>
__device__ __forceinline__ void seam_reassign_body(const int* __restrict__ boxes, int ci, int cj, int w,
                                                   Own* __restrict__ owner, const int* __restrict__ lohi,
                                                   const int* __restrict__ chosen, const int* __restrict__ contested,
                                                   int line) {
    SeamGeom g;
    if (!seam_geom(boxes, ci, cj, g)) return;
    if (line >= g.lines || *gptr(contested) == 0) return;  // wave-uniform
    const int ch = gptr(chosen)[line];
    if (ch < 0) return;
    const int lo = gptr(lohi)[2 * line], hi = gptr(lohi)[2 * line + 1];
    const int t = lo + (threadIdx.x & 63);
    if (t > hi) return;
    const size_t p = seam_at(g.rowScan, g.ox0, g.oy0, w, line, t);
    const int o = gptr(owner)[p];
    if (o != ci && o != cj) return;
    gptr(owner)[p] = static_cast<Own>(((t < ch) == g.iLow) ? ci : cj);
}

template <class Img, typename M>
__global__ void __launch_bounds__(kDpThreads) seam_dp_kernel(const int* __restrict__ boxes, int ci, int cj,
                                                             const M* __restrict__ costs,
                                                             const int* __restrict__ lohi, int* __restrict__ prevl,
                                                             int* __restrict__ chosen, uint8_t* __restrict__ back,
                                                             const int* __restrict__ contested) {
    seam_dp_body<Img, M>(contested, boxes, ci, cj, costs, lohi, prevl, chosen, back);
}

template <int Dummy = 0>  // a template, so the header defines it once per program
__global__ void __launch_bounds__(256) seam_reassign_kernel(const int* __restrict__ boxes, int ci, int cj, int w,
                                                            int32_t* __restrict__ owner,
                                                            const int* __restrict__ lohi,
                                                            const int* __restrict__ chosen,
                                                            const int* __restrict__ contested) {
    const int line = blockIdx.x * kScanLinesPerBlock + (threadIdx.x >> 6);
    seam_reassign_body(boxes, ci, cj, w, owner, lohi, chosen, contested, line);
}

// One pair of one chunk per table entry (batched form: every chunk's k-th
// pair in one launch).
template <class Img>
struct SeamWork {
    Img imgI, imgJ;
    const uint8_t* covI;
    const uint8_t* covJ;
    int covShift;      // coverage layout: < 0 row-major, else sheared (seam_scan_body)
    const int* boxes;  // the chunk's contributor boxes
    int ci, cj, w;
    int8_t* owner;  // the float render's owner map (contributor index, -1: none)
    int* lohi;
    int* prevl;
    int* chosen;
    uint8_t* back;
    void* costs;     // 64 cost values (of the DP's arithmetic type) per line
    int* contested;  // contested-line count of this pair, cleared before the render
};

template <class Img, typename M>
__global__ void __launch_bounds__(256) seam_scan_table_kernel(const SeamWork<Img>* __restrict__ work) {
    const SeamWork<Img> wk = work[blockIdx.y];
    const int line = blockIdx.x * kScanLinesPerBlock + (threadIdx.x >> 6);
    seam_scan_body<Img, M, int8_t>(wk.imgI, wk.imgJ, wk.covI, wk.covJ, wk.covShift, wk.boxes, wk.ci, wk.cj, wk.w, wk.owner,
                                   wk.lohi,
                           wk.chosen, wk.contested, static_cast<M*>(wk.costs), line);
}

template <class Img, typename M>
__global__ void __launch_bounds__(kDpThreads) seam_dp_table_kernel(const SeamWork<Img>* __restrict__ work) {
    const SeamWork<Img> wk = work[blockIdx.x];
    seam_dp_body<Img, M>(wk.contested, wk.boxes, wk.ci, wk.cj, static_cast<const M*>(wk.costs), wk.lohi, wk.prevl,
                         wk.chosen, wk.back);
}

template <class Img>
__global__ void __launch_bounds__(256) seam_reassign_table_kernel(const SeamWork<Img>* __restrict__ work) {
    const SeamWork<Img> wk = work[blockIdx.y];
    const int line = blockIdx.x * kScanLinesPerBlock + (threadIdx.x >> 6);
    seam_reassign_body(wk.boxes, wk.ci, wk.cj, wk.w, wk.owner, wk.lohi, wk.chosen, wk.contested, line);
}

}  // namespace i360
