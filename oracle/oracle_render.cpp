// Oracle for the float chunk render (TEST INFRASTRUCTURE ONLY; see
// oracle.cpp's header for the rules that apply to this directory).
//
// The float render is the reference's per-chunk render
// (WindowsApp.Core/SourceFiles/RenderExecutor.cpp:70-140: warp every
// contributor, apply its gain, SeamBlend::BlendChunkContributors) carried
// out on float RGB in [0,1] instead of RGB48:
//   * warp: the bilinear rule of WarpPerspectiveKernels.h:16-43 and the
//     validity rule of WarpPerspective_Scalar.cpp:117-123, on float texels;
//     the output carries an explicit coverage bit (valid sample) instead of
//     the RGB48 "(0,0,0) means no data" convention;
//   * gain: v * gain, saturated to [0,1] (GainColorOps.cpp:16-26 with 1.0
//     standing for 65535);
//   * ownership: the two-pass chamfer of SeamBlend.cpp:206-242 on the
//     coverage bits (always in binary32: it is a decision procedure on
//     binary data), argmax of SeamBlend.cpp:244-261 and the pairwise DP seam
//     of SeamBlend.cpp:293-445 with colour costs in [0,1] units and the
//     smoothness weight rescaled to 2/65535 per pixel;
//   * compose: SeamBlend.cpp:447-546 in the scalar type T, output clamped to
//     [0,1] (no u16 rounding), zero where no contributor covers.
// Loops whose iterations write disjoint outputs with the same arithmetic
// run on OpenMP threads (rows of a plane, pixels, contributors' chamfer
// planes); every value is computed exactly as in a serial run, and the
// accumulation over contributors keeps its order.
// T = double is the FP64 ground truth (inverse homography and coordinate
// map in binary64 too); T = float with a binary32 inverse homography is the
// plain FP32 form.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

extern "C" void orc_coverage_distance(const uint8_t* cov, int w, int h, float* d);
extern "C" void orc_ownership(const float* const* dist, int c, int w, int h, int32_t* owner);
extern "C" int orc_cover_box(const uint8_t* cov, int w, int h, int* box);
extern "C" int orc_clamp_bands(int w, int h, int numBands);

namespace {

inline size_t Px(int x, int y, int w) { return static_cast<size_t>(y) * w + x; }

// Planted mistakes, for the tests that show a pin has teeth (0 = none):
// 1 swaps the bilinear weights w10 and w01; 2 applies the gain after the
// clamp.  Set only by orc_render_plant, from those tests.
int g_plant = 0;

template <typename T>
void Blur5T(const std::vector<T>& src, int w, int h, int ch, std::vector<T>& dst) {
    const T k0 = T(1) / T(16), k1 = T(4) / T(16), k2 = T(6) / T(16);
    std::vector<T> tmp(src.size());
    dst.assign(src.size(), T(0));
    auto cx = [w](int x) { return x < 0 ? 0 : (x > w - 1 ? w - 1 : x); };
    auto cy = [h](int y) { return y < 0 ? 0 : (y > h - 1 ? h - 1 : y); };
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < ch; ++c) {
                const T* row = src.data() + static_cast<size_t>(y) * w * ch;
                T a = k0 * row[cx(x - 2) * ch + c], b = k1 * row[cx(x - 1) * ch + c], m = k2 * row[cx(x) * ch + c];
                T d = k1 * row[cx(x + 1) * ch + c], e = k0 * row[cx(x + 2) * ch + c];
                tmp[Px(x, y, w) * ch + c] = (((a + b) + m) + d) + e;
            }
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < ch; ++c) {
                T a = k0 * tmp[Px(x, cy(y - 2), w) * ch + c], b = k1 * tmp[Px(x, cy(y - 1), w) * ch + c];
                T m = k2 * tmp[Px(x, cy(y), w) * ch + c], d = k1 * tmp[Px(x, cy(y + 1), w) * ch + c];
                T e = k0 * tmp[Px(x, cy(y + 2), w) * ch + c];
                dst[Px(x, y, w) * ch + c] = (((a + b) + m) + d) + e;
            }
}

template <typename T>
void ReduceT(const std::vector<T>& src, int w, int h, int ch, std::vector<T>& dst, int& ow, int& oh) {
    std::vector<T> bl;
    Blur5T(src, w, h, ch, bl);
    ow = (w + 1) / 2;
    oh = (h + 1) / 2;
    dst.assign(static_cast<size_t>(ow) * oh * ch, T(0));
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x)
            for (int c = 0; c < ch; ++c)
                dst[Px(x, y, ow) * ch + c] = bl[Px(std::min(2 * x, w - 1), std::min(2 * y, h - 1), w) * ch + c];
}

template <typename T>
void ExpandT(const std::vector<T>& src, int w, int h, int ch, int ow, int oh, std::vector<T>& dst) {
    dst.assign(static_cast<size_t>(ow) * oh * ch, T(0));
    T sx = (w > 1) ? T(w - 1) / T(std::max(1, ow - 1)) : T(0);
    T sy = (h > 1) ? T(h - 1) / T(std::max(1, oh - 1)) : T(0);
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < oh; ++y) {
        T fy = sy * T(y);
        int y0 = std::min(static_cast<int>(fy), h - 1), y1 = std::min(y0 + 1, h - 1);
        T wy = fy - T(y0);
        for (int x = 0; x < ow; ++x) {
            T fx = sx * T(x);
            int x0 = std::min(static_cast<int>(fx), w - 1), x1 = std::min(x0 + 1, w - 1);
            T wx = fx - T(x0);
            for (int c = 0; c < ch; ++c) {
                T v00 = src[Px(x0, y0, w) * ch + c], v10 = src[Px(x1, y0, w) * ch + c];
                T v01 = src[Px(x0, y1, w) * ch + c], v11 = src[Px(x1, y1, w) * ch + c];
                T omx = T(1) - wx, omy = T(1) - wy;
                T top = omx * v00 + wx * v10, bot = omx * v01 + wx * v11;
                dst[Px(x, y, ow) * ch + c] = omy * top + wy * bot;
            }
        }
    }
}

template <typename T>
struct LevelT { std::vector<T> d; int w = 0, h = 0; };

template <typename T>
void GaussPyrT(const std::vector<T>& base, int w, int h, int ch, int L, std::vector<LevelT<T>>& out) {
    out.assign(L, LevelT<T>{});
    out[0].d = base; out[0].w = w; out[0].h = h;
    for (int l = 1; l < L; ++l) ReduceT(out[l - 1].d, out[l - 1].w, out[l - 1].h, ch, out[l].d, out[l].w, out[l].h);
}

// Warp one contributor into the chunk (float RGB + coverage bit).
template <typename T>
void WarpF(const float* src, int sw, int sh, const T* m, float gain, int x0, int y0, int w, int h,
           std::vector<T>& rgb, std::vector<uint8_t>& cov) {
    size_t n = static_cast<size_t>(w) * h;
    rgb.assign(n * 3, T(0));
    cov.assign(n, 0);
    const bool useGain = std::fabs(gain - 1.0f) >= 1e-6f;
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        T Y = T(y0 + y);
        for (int x = 0; x < w; ++x) {
            T X = T(x0 + x);
            T u = (m[0] * X + m[1] * Y) + m[2];
            T v = (m[3] * X + m[4] * Y) + m[5];
            T q = (m[6] * X + m[7] * Y) + m[8];
            if (std::fabs(q) < T(1e-10)) continue;
            u = u / q;
            v = v / q;
            T fu = std::floor(u), fv = std::floor(v);
            if (!(fu >= T(0) && fv >= T(0) && fu <= T(sw - 2) && fv <= T(sh - 2))) continue;
            int sx0 = static_cast<int>(fu), sy0 = static_cast<int>(fv);
            T fx = u - T(sx0), fy = v - T(sy0);
            T omx = T(1) - fx, omy = T(1) - fy;
            T w00 = omx * omy, w10 = fx * omy, w01 = omx * fy, w11 = fx * fy;
            if (g_plant == 1) std::swap(w10, w01);
            const float* r0 = src + (static_cast<size_t>(sy0) * sw + sx0) * 3;
            const float* r1 = r0 + static_cast<size_t>(sw) * 3;
            size_t p = Px(x, y, w);
            cov[p] = 1;
            for (int c = 0; c < 3; ++c) {
                T val = ((w00 * T(r0[c]) + w10 * T(r0[3 + c])) + w01 * T(r1[c])) + w11 * T(r1[3 + c]);
                if (g_plant == 2) {
                    val = val < T(0) ? T(0) : (T(1) < val ? T(1) : val);
                    if (useGain) val = val * T(gain);
                } else {
                    if (useGain) val = val * T(gain);
                    val = val < T(0) ? T(0) : (T(1) < val ? T(1) : val);
                }
                rgb[p * 3 + c] = val;
            }
        }
    }
}

// SeamBlend.cpp:293-445 on float-valued images (costs in [0,1] units).
template <typename T>
void SeamDPT(const T* imgI, const T* imgJ, const uint8_t* covI, const uint8_t* covJ, int w, int ci, int cj,
             const int* bI, const int* bJ, const int* ov, int32_t* owner) {
    const int ox0 = ov[0], oy0 = ov[1], ox1 = ov[2], oy1 = ov[3];
    const int bw = ox1 - ox0 + 1, bh = oy1 - oy0 + 1;
    const bool rowScan = bw >= bh;
    float cI = rowScan ? float(bI[0] + bI[2]) * 0.5f : float(bI[1] + bI[3]) * 0.5f;
    float cJ = rowScan ? float(bJ[0] + bJ[2]) * 0.5f : float(bJ[1] + bJ[3]) * 0.5f;
    const bool iLow = cI <= cJ;
    const T kSmooth = T(2) / T(65535);
    const int kRadius = 24;
    const int lines = rowScan ? bh : bw, len = rowScan ? bw : bh;
    auto at = [&](int line, int t) -> size_t {
        return rowScan ? Px(ox0 + t, oy0 + line, w) : Px(ox0 + line, oy0 + t, w);
    };
    std::vector<int> lo(lines, 1), hi(lines, 0);
    for (int ln = 0; ln < lines; ++ln) {
        int cross = -1, first = -1;
        for (int t = 0; t < len; ++t) {
            int o = owner[at(ln, t)];
            if (o != ci && o != cj) continue;
            if (first < 0) { first = o; continue; }
            if (o != first) { cross = t; break; }
        }
        if (cross < 0) continue;
        int a = std::max(0, cross - kRadius), b = std::min(len - 1, cross + kRadius);
        for (int t = a; t <= b; ++t) {
            size_t p = at(ln, t);
            if (covI[p] && covJ[p]) {
                if (hi[ln] < lo[ln]) { lo[ln] = t; hi[ln] = t; }
                else { lo[ln] = std::min(lo[ln], t); hi[ln] = std::max(hi[ln], t); }
            }
        }
    }
    std::vector<std::vector<T>> dp(lines);
    std::vector<std::vector<int>> bk(lines);
    std::vector<int> prevLine(lines, -1);
    int last = -1;
    for (int ln = 0; ln < lines; ++ln) {
        if (hi[ln] < lo[ln]) continue;
        int lw = hi[ln] - lo[ln] + 1;
        dp[ln].assign(lw, T(0));
        bk[ln].assign(lw, -1);
        for (int k = 0; k < lw; ++k) {
            int t = lo[ln] + k;
            size_t p = at(ln, t);
            T cost = T(0);
            for (int c = 0; c < 3; ++c) cost = cost + std::fabs(imgI[p * 3 + c] - imgJ[p * 3 + c]);
            if (last < 0) { dp[ln][k] = cost; continue; }
            T best = std::numeric_limits<T>::max();
            int bp = -1;
            for (int pk = 0; pk < static_cast<int>(dp[last].size()); ++pk) {
                T cand = dp[last][pk] + kSmooth * T(std::abs(t - (lo[last] + pk)));
                if (cand < best) { best = cand; bp = pk; }
            }
            dp[ln][k] = cost + best;
            bk[ln][k] = bp;
        }
        prevLine[ln] = last;
        last = ln;
    }
    if (last < 0) return;
    std::vector<int> chosen(lines, -1);
    int ln = last, bestK = 0;
    {
        T best = std::numeric_limits<T>::max();
        for (size_t k = 0; k < dp[ln].size(); ++k)
            if (dp[ln][k] < best) { best = dp[ln][k]; bestK = static_cast<int>(k); }
    }
    while (ln >= 0) {
        chosen[ln] = lo[ln] + bestK;
        int pl = prevLine[ln], pk = bk[ln][bestK];
        ln = pl;
        bestK = pk;
    }
    for (int l = 0; l < lines; ++l) {
        if (chosen[l] < 0) continue;
        for (int t = lo[l]; t <= hi[l]; ++t) {
            size_t p = at(l, t);
            if (owner[p] != ci && owner[p] != cj) continue;
            owner[p] = ((t < chosen[l]) == iLow) ? ci : cj;
        }
    }
}

// Owner map for float-valued warped contributors (chamfer -> argmax -> seams).
template <typename T>
void OwnerMapT(const std::vector<std::vector<T>>& rgb, const std::vector<std::vector<uint8_t>>& cov, int w, int h,
               std::vector<int32_t>& owner) {
    int c = static_cast<int>(rgb.size());
    size_t n = static_cast<size_t>(w) * h;
    std::vector<std::vector<float>> dist(c, std::vector<float>(n));
    std::vector<const float*> dp(c);
    std::vector<std::array<int, 4>> box(c);
    std::vector<int> valid(c);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < c; ++i) {
        orc_coverage_distance(cov[i].data(), w, h, dist[i].data());
        dp[i] = dist[i].data();
        valid[i] = orc_cover_box(cov[i].data(), w, h, box[i].data());
    }
    owner.assign(n, -1);
    orc_ownership(dp.data(), c, w, h, owner.data());
    for (int i = 0; i < c; ++i) {
        if (!valid[i]) continue;
        for (int j = i + 1; j < c; ++j) {
            if (!valid[j]) continue;
            int o[4] = {std::max(box[i][0], box[j][0]), std::max(box[i][1], box[j][1]),
                        std::min(box[i][2], box[j][2]), std::min(box[i][3], box[j][3])};
            if (o[0] > o[2] || o[1] > o[3]) continue;
            SeamDPT<T>(rgb[i].data(), rgb[j].data(), cov[i].data(), cov[j].data(), w, i, j, box[i].data(),
                       box[j].data(), o, owner.data());
        }
    }
}

template <typename T>
void ComposeT(const std::vector<std::vector<T>>& rgb, const std::vector<std::vector<uint8_t>>& cov,
              const int32_t* owner, int w, int h, int numBands, float* out) {
    int c = static_cast<int>(rgb.size());
    size_t n = static_cast<size_t>(w) * h;
    std::memset(out, 0, n * 3 * sizeof(float));
    int L = orc_clamp_bands(w, h, numBands);
    std::vector<bool> owns(c, false);
    for (size_t p = 0; p < n; ++p) if (owner[p] >= 0) owns[owner[p]] = true;
    std::vector<LevelT<T>> band;
    bool init = false;
    for (int i = 0; i < c; ++i) {
        if (!owns[i]) continue;
        std::vector<T> wm(n), num(n * 3), den(n);
        #pragma omp parallel for schedule(static)
        for (size_t p = 0; p < n; ++p) {
            wm[p] = owner[p] == i ? T(1) : T(0);
            T m = cov[i][p] ? T(1) : T(0);
            den[p] = m;
            for (int k = 0; k < 3; ++k) num[p * 3 + k] = rgb[i][p * 3 + k] * m;
        }
        std::vector<LevelT<T>> wp, np, dpy;
        GaussPyrT(wm, w, h, 1, L, wp);
        GaussPyrT(num, w, h, 3, L, np);
        GaussPyrT(den, w, h, 1, L, dpy);
        std::vector<LevelT<T>> g(L);
        for (int l = 0; l < L; ++l) {
            g[l].w = np[l].w; g[l].h = np[l].h;
            size_t m = static_cast<size_t>(g[l].w) * g[l].h;
            g[l].d.resize(m * 3);
            #pragma omp parallel for schedule(static)
            for (size_t p = 0; p < m; ++p) {
                T dv = dpy[l].d[p];
                T inv = dv > T(1e-6) ? T(1) / dv : T(0);
                for (int k = 0; k < 3; ++k) g[l].d[p * 3 + k] = np[l].d[p * 3 + k] * inv;
            }
        }
        if (!init) {
            band.assign(L, LevelT<T>{});
            for (int l = 0; l < L; ++l) { band[l].w = g[l].w; band[l].h = g[l].h; band[l].d.assign(g[l].d.size(), T(0)); }
            init = true;
        }
        for (int l = 0; l < L; ++l) {
            std::vector<T> up;
            if (l + 1 < L) ExpandT(g[l + 1].d, g[l + 1].w, g[l + 1].h, 3, g[l].w, g[l].h, up);
            size_t m = static_cast<size_t>(band[l].w) * band[l].h;
            #pragma omp parallel for schedule(static)
            for (size_t p = 0; p < m; ++p) {
                T wt = wp[l].d[p];
                for (int k = 0; k < 3; ++k) {
                    T lap = (l + 1 < L) ? g[l].d[p * 3 + k] - up[p * 3 + k] : g[l].d[p * 3 + k];
                    band[l].d[p * 3 + k] = band[l].d[p * 3 + k] + wt * lap;
                }
            }
        }
    }
    if (!init) return;
    std::vector<T> rec = band[L - 1].d;
    int rw = band[L - 1].w, rh = band[L - 1].h;
    for (int l = L - 2; l >= 0; --l) {
        std::vector<T> up;
        ExpandT(rec, rw, rh, 3, band[l].w, band[l].h, up);
        rec.resize(up.size());
        #pragma omp parallel for schedule(static)
        for (size_t p = 0; p < up.size(); ++p) rec[p] = band[l].d[p] + up[p];
        rw = band[l].w; rh = band[l].h;
    }
    #pragma omp parallel for schedule(static)
    for (size_t p = 0; p < n; ++p)
        for (int k = 0; k < 3; ++k) {
            T v = rec[p * 3 + k];
            v = v < T(1) ? v : T(1);
            v = T(0) < v ? v : T(0);
            out[p * 3 + k] = owner[p] >= 0 ? static_cast<float>(v) : 0.0f;
        }
}

template <typename T>
void RenderChunkT(const float* const* srcs, const int* sw, const int* sh, const T* hinv, const float* gains, int c,
                  int x0, int y0, int w, int h, int numBands, float* out, int32_t* ownerOut, const int32_t* ownerIn) {
    std::vector<std::vector<T>> rgb(c);
    std::vector<std::vector<uint8_t>> cov(c);
    for (int i = 0; i < c; ++i) WarpF<T>(srcs[i], sw[i], sh[i], hinv + 9 * i, gains[i], x0, y0, w, h, rgb[i], cov[i]);
    std::vector<int32_t> owner;
    if (ownerIn) owner.assign(ownerIn, ownerIn + static_cast<size_t>(w) * h);
    else OwnerMapT<T>(rgb, cov, w, h, owner);
    if (ownerOut) std::memcpy(ownerOut, owner.data(), owner.size() * sizeof(int32_t));
    ComposeT<T>(rgb, cov, owner.data(), w, h, numBands, out);
}

}  // namespace

extern "C" {

// srcs: c float RGB frames (the stored FP16/FP32 values widened to float);
// hinv: 9 values per contributor, canvas -> source, row-major.
// ownerIn: optional owner map to use instead of computing one.
void orc_render_chunk_f64(const float* const* srcs, const int* sw, const int* sh, const double* hinv,
                          const float* gains, int c, int x0, int y0, int w, int h, int numBands, float* out,
                          int32_t* ownerOut, const int32_t* ownerIn) {
    RenderChunkT<double>(srcs, sw, sh, hinv, gains, c, x0, y0, w, h, numBands, out, ownerOut, ownerIn);
}

void orc_render_chunk_f32(const float* const* srcs, const int* sw, const int* sh, const float* hinv,
                          const float* gains, int c, int x0, int y0, int w, int h, int numBands, float* out,
                          int32_t* ownerOut, const int32_t* ownerIn) {
    RenderChunkT<float>(srcs, sw, sh, hinv, gains, c, x0, y0, w, h, numBands, out, ownerOut, ownerIn);
}

// One contributor's warp into a chunk (the render's first step): float RGB
// (w*h*3) and the coverage bit (w*h) of every chunk pixel.
void orc_warp_float_f64(const float* src, int sw, int sh, const double* hinv, float gain, int x0, int y0, int w,
                        int h, double* rgb, uint8_t* cov) {
    std::vector<double> r;
    std::vector<uint8_t> c;
    WarpF<double>(src, sw, sh, hinv, gain, x0, y0, w, h, r, c);
    std::memcpy(rgb, r.data(), r.size() * sizeof(double));
    std::memcpy(cov, c.data(), c.size());
}

void orc_warp_float_f32(const float* src, int sw, int sh, const float* hinv, float gain, int x0, int y0, int w,
                        int h, float* rgb, uint8_t* cov) {
    std::vector<float> r;
    std::vector<uint8_t> c;
    WarpF<float>(src, sw, sh, hinv, gain, x0, y0, w, h, r, c);
    std::memcpy(rgb, r.data(), r.size() * sizeof(float));
    std::memcpy(cov, c.data(), c.size());
}

// Test hook: plants mistake `which` (see g_plant) in the float warp; 0 removes it.
void orc_render_plant(int which) { g_plant = which; }

// FP64 adjugate inverse of a forward homography (local -> canvas), the
// binary64 counterpart of HomographyMath.cpp:94-116.
int orc_invert_homography_f64(const double* m, double* inv) {
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], i = m[7], j = m[8];
    double c0 = e * j - f * i, c1 = d * j - f * g, c2 = d * i - e * g;
    double det = (a * c0 - b * c1) + c * c2;
    if (std::fabs(det) < 1e-300) return 0;
    double s = 1.0 / det;
    inv[0] = c0 * s; inv[1] = (c * i - b * j) * s; inv[2] = (b * f - c * e) * s;
    inv[3] = (f * g - d * j) * s; inv[4] = (a * j - c * g) * s; inv[5] = (c * d - a * f) * s;
    inv[6] = c2 * s; inv[7] = (b * g - a * i) * s; inv[8] = (a * e - b * d) * s;
    return 1;
}

}  // extern "C"
