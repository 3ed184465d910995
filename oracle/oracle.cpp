// Oracle: a scalar CPU restatement of the reference CPU backend's hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
// calls this file; only the test suite, the smoke check and the benchmark's
// CPU-baseline leg may use it, and only as the checker.
//
// Every function restates the behaviour of the reference's Scalar tier
// (file:line cited per function, paths relative to the reference root) in
// code written for this repository.  It is compiled with
// -ffp-contract=off and without -ffast-math, so every float operation is a
// single IEEE-754 binary32 operation in source order -- the same sequence
// the reference Scalar tier executes (the Scalar TUs are built without
// -mfma, so nothing there is contracted either).
//
// Parity pins (see tests/test_oracle_*.py): each function is checked against
// golden vectors written by the compiled reference sources
// (tests/golden/), against oracle/_ref when that library is present, and
// against closed forms / brute force on tiny inputs.
//
// FP64 render functions at the bottom ("ground truth" for PSNR) follow the
// same algorithm with every intermediate in binary64 and no u16 rounding.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

inline uint16_t SatU16(float v) {
    // (std::min)((std::max)(v, 0.0f), 65535.0f) then truncation
    // (MedianStack/Warp/Demosaic all use this form, e.g.
    // WindowsApp.Core/HeaderFiles/WarpPerspectiveKernels.h:41).
    float a = (v < 0.0f) ? 0.0f : v;
    float b = (65535.0f < a) ? 65535.0f : a;
    return static_cast<uint16_t>(b);
}

// ---------------------------------------------------------------- median --
// WindowsApp.Core/HeaderFiles/MedianStackKernels.h:19-55 and
// WindowsApp.Core/SourceFiles/MedianStack_Scalar.cpp:8-24.
uint16_t MedianOne(uint16_t* v, int count, float sigmaThreshold) {
    if (count > 3) {
        float sum = 0.0f, sumSq = 0.0f;
        for (int i = 0; i < count; ++i) {
            float f = static_cast<float>(v[i]);
            sum = sum + f;
            float sq = f * f;
            sumSq = sumSq + sq;
        }
        float mean = sum / static_cast<float>(count);
        float msq = mean * mean;
        float variance = sumSq / static_cast<float>(count) - msq;
        float sigma = std::sqrt(variance > 0.0f ? variance : 0.0f);
        float ts = sigmaThreshold * sigma;
        float lower = mean - ts;
        float upper = mean + ts;
        int kept = 0;
        for (int i = 0; i < count; ++i) {
            float f = static_cast<float>(v[i]);
            if (f >= lower && f <= upper) v[kept++] = v[i];
        }
        if (kept > 0) count = kept;
    }
    std::sort(v, v + count);
    return v[count / 2];  // upper median
}

// ------------------------------------------------------------ demosaic --
// WindowsApp.Core/HeaderFiles/BayerDemosaicKernels.h:10-13 (LibRaw COLOR()).
inline int CfaColor(int row, int col, uint32_t filters) {
    return static_cast<int>((filters >> ((((row << 1) & 14) | (col & 1)) << 1)) & 3u);
}

// Shared body of the u16 and f32 demosaic entry points.
// BayerDemosaic_Scalar.cpp:9-50 with BayerDemosaicKernels.h:25-75.
template <bool kFloatOut, typename OutT>
void Demosaic(const uint16_t* cfaIn, int w, int h, uint16_t black, const float* camMul,
              const float* rgbCam, uint32_t filters, OutT* out) {
    const size_t n = static_cast<size_t>(w) * h;
    std::vector<uint16_t> cfa(n);
    // (1) black level, saturating (BayerDemosaic_Scalar.cpp:19-22)
    for (size_t i = 0; i < n; ++i) cfa[i] = cfaIn[i] > black ? static_cast<uint16_t>(cfaIn[i] - black) : 0;
    // (2) white balance per CFA colour, channel 3 = second green (:25-35)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = static_cast<size_t>(y) * w + x;
            cfa[i] = SatU16(static_cast<float>(cfa[i]) * camMul[CfaColor(y, x, filters)]);
        }
    // (3) same-colour 3x3 bilinear, own sample exact (BayerDemosaicKernels.h:25-60)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            float acc[3] = {0.0f, 0.0f, 0.0f};
            int cnt[3] = {0, 0, 0};
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int nx = x + dx, ny = y + dy;
                    if (nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
                    int ch = CfaColor(ny, nx, filters);
                    int s = ch == 3 ? 1 : ch;
                    acc[s] = acc[s] + static_cast<float>(cfa[static_cast<size_t>(ny) * w + nx]);
                    cnt[s] += 1;
                }
            int mine = CfaColor(y, x, filters);
            int ms = mine == 3 ? 1 : mine;
            for (int c = 0; c < 3; ++c)
                if (c != ms && cnt[c] != 0) acc[c] = acc[c] / static_cast<float>(cnt[c]);
            acc[ms] = static_cast<float>(cfa[static_cast<size_t>(y) * w + x]);
            uint16_t q[3] = {SatU16(acc[0]), SatU16(acc[1]), SatU16(acc[2])};
            // (4) 3x4 colour matrix, left-to-right sum (BayerDemosaicKernels.h:64-75)
            float r = q[0], g = q[1], b = q[2];
            for (int row = 0; row < 3; ++row) {
                const float* m = rgbCam + row * 4;
                float t0 = m[0] * r;
                float t1 = m[1] * g;
                float t2 = m[2] * b;
                float val = ((t0 + t1) + t2) + m[3];
                size_t o = (static_cast<size_t>(y) * w + x) * 3 + row;
                if constexpr (kFloatOut) {
                    float a = (val < 0.0f) ? 0.0f : val;
                    out[o] = (65535.0f < a) ? 65535.0f : a;
                } else {
                    out[o] = SatU16(val);
                }
            }
        }
    }
}

// ---------------------------------------------------------- seam blend --
inline size_t Px(int x, int y, int w) { return static_cast<size_t>(y) * w + x; }

// SeamBlend.cpp:15-56: separable [1,4,6,4,1]/16, replicate border, the
// horizontal pass first, each 5-term sum evaluated left to right.
void Blur5(const std::vector<float>& src, int w, int h, int ch, std::vector<float>& dst) {
    const float k0 = 1.0f / 16.0f, k1 = 4.0f / 16.0f, k2 = 6.0f / 16.0f;
    std::vector<float> tmp(src.size());
    dst.assign(src.size(), 0.0f);
    auto cx = [w](int x) { return x < 0 ? 0 : (x > w - 1 ? w - 1 : x); };
    auto cy = [h](int y) { return y < 0 ? 0 : (y > h - 1 ? h - 1 : y); };
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < ch; ++c) {
                const float* row = src.data() + static_cast<size_t>(y) * w * ch;
                float a = k0 * row[cx(x - 2) * ch + c];
                float b = k1 * row[cx(x - 1) * ch + c];
                float m = k2 * row[cx(x) * ch + c];
                float d = k1 * row[cx(x + 1) * ch + c];
                float e = k0 * row[cx(x + 2) * ch + c];
                tmp[Px(x, y, w) * ch + c] = (((a + b) + m) + d) + e;
            }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < ch; ++c) {
                float a = k0 * tmp[Px(x, cy(y - 2), w) * ch + c];
                float b = k1 * tmp[Px(x, cy(y - 1), w) * ch + c];
                float m = k2 * tmp[Px(x, cy(y), w) * ch + c];
                float d = k1 * tmp[Px(x, cy(y + 1), w) * ch + c];
                float e = k0 * tmp[Px(x, cy(y + 2), w) * ch + c];
                dst[Px(x, y, w) * ch + c] = (((a + b) + m) + d) + e;
            }
}

// SeamBlend.cpp:60-79: blur, then keep (min(2x,w-1), min(2y,h-1)).
void Reduce(const std::vector<float>& src, int w, int h, int ch, std::vector<float>& dst, int& ow, int& oh) {
    std::vector<float> bl;
    Blur5(src, w, h, ch, bl);
    ow = (w + 1) / 2;
    oh = (h + 1) / 2;
    dst.assign(static_cast<size_t>(ow) * oh * ch, 0.0f);
    for (int y = 0; y < oh; ++y) {
        int sy = std::min(2 * y, h - 1);
        for (int x = 0; x < ow; ++x) {
            int sx = std::min(2 * x, w - 1);
            for (int c = 0; c < ch; ++c) dst[Px(x, y, ow) * ch + c] = bl[Px(sx, sy, w) * ch + c];
        }
    }
}

// SeamBlend.cpp:85-115: align-corners bilinear resize.
void Expand(const std::vector<float>& src, int w, int h, int ch, int ow, int oh, std::vector<float>& dst) {
    dst.assign(static_cast<size_t>(ow) * oh * ch, 0.0f);
    float sx = (w > 1) ? static_cast<float>(w - 1) / static_cast<float>(std::max(1, ow - 1)) : 0.0f;
    float sy = (h > 1) ? static_cast<float>(h - 1) / static_cast<float>(std::max(1, oh - 1)) : 0.0f;
    for (int y = 0; y < oh; ++y) {
        float fy = sy * static_cast<float>(y);
        int y0 = std::min(static_cast<int>(fy), h - 1);
        int y1 = std::min(y0 + 1, h - 1);
        float wy = fy - static_cast<float>(y0);
        for (int x = 0; x < ow; ++x) {
            float fx = sx * static_cast<float>(x);
            int x0 = std::min(static_cast<int>(fx), w - 1);
            int x1 = std::min(x0 + 1, w - 1);
            float wx = fx - static_cast<float>(x0);
            for (int c = 0; c < ch; ++c) {
                float v00 = src[Px(x0, y0, w) * ch + c], v10 = src[Px(x1, y0, w) * ch + c];
                float v01 = src[Px(x0, y1, w) * ch + c], v11 = src[Px(x1, y1, w) * ch + c];
                float omx = 1.0f - wx, omy = 1.0f - wy;
                float top = omx * v00 + wx * v10;
                float bot = omx * v01 + wx * v11;
                dst[Px(x, y, ow) * ch + c] = omy * top + wy * bot;
            }
        }
    }
}

struct Level { std::vector<float> d; int w = 0, h = 0; };

void GaussPyr(const std::vector<float>& base, int w, int h, int ch, int L, std::vector<Level>& out) {
    out.assign(L, Level{});
    out[0].d = base; out[0].w = w; out[0].h = h;
    for (int l = 1; l < L; ++l) Reduce(out[l - 1].d, out[l - 1].w, out[l - 1].h, ch, out[l].d, out[l].w, out[l].h);
}

// SeamBlend.cpp:150-182 (normalised convolution; 1/den then multiply).
void HoleFilledPyr(const uint16_t* img, const std::vector<float>& cov, int w, int h, int L, std::vector<Level>& out) {
    size_t n = static_cast<size_t>(w) * h;
    std::vector<float> num(n * 3), den(n);
    for (size_t p = 0; p < n; ++p) {
        den[p] = cov[p];
        for (int c = 0; c < 3; ++c) num[p * 3 + c] = static_cast<float>(img[p * 3 + c]) * cov[p];
    }
    std::vector<Level> np, dp;
    GaussPyr(num, w, h, 3, L, np);
    GaussPyr(den, w, h, 1, L, dp);
    out.assign(L, Level{});
    for (int l = 0; l < L; ++l) {
        out[l].w = np[l].w; out[l].h = np[l].h;
        size_t m = static_cast<size_t>(np[l].w) * np[l].h;
        out[l].d.assign(m * 3, 0.0f);
        for (size_t p = 0; p < m; ++p) {
            float dv = dp[l].d[p];
            float inv = (dv > 1e-6f) ? (1.0f / dv) : 0.0f;
            for (int c = 0; c < 3; ++c) out[l].d[p * 3 + c] = np[l].d[p * 3 + c] * inv;
        }
    }
}

// SeamBlend.cpp:184-201.
void LaplacianPyr(const std::vector<Level>& g, std::vector<Level>& out) {
    int L = static_cast<int>(g.size());
    out.assign(L, Level{});
    for (int l = 0; l + 1 < L; ++l) {
        std::vector<float> up;
        Expand(g[l + 1].d, g[l + 1].w, g[l + 1].h, 3, g[l].w, g[l].h, up);
        out[l].w = g[l].w; out[l].h = g[l].h;
        out[l].d.resize(up.size());
        for (size_t p = 0; p < up.size(); ++p) out[l].d[p] = g[l].d[p] - up[p];
    }
    out[L - 1] = g[L - 1];
}

struct Box { int x0 = 0, y0 = 0, x1 = -1, y1 = -1; bool valid = false; };

Box CoverBox(const uint8_t* cov, int w, int h) {  // SeamBlend.cpp:263-280
    Box b;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            if (!cov[Px(x, y, w)]) continue;
            if (!b.valid) { b.x0 = b.x1 = x; b.y0 = b.y1 = y; b.valid = true; }
            else { b.x0 = std::min(b.x0, x); b.x1 = std::max(b.x1, x); b.y0 = std::min(b.y0, y); b.y1 = std::max(b.y1, y); }
        }
    return b;
}

bool Intersect(const Box& a, const Box& b, Box& o) {  // SeamBlend.cpp:282-291
    if (!a.valid || !b.valid) { o = Box{}; return false; }
    o.x0 = std::max(a.x0, b.x0); o.y0 = std::max(a.y0, b.y0);
    o.x1 = std::min(a.x1, b.x1); o.y1 = std::min(a.y1, b.y1);
    o.valid = o.x0 <= o.x1 && o.y0 <= o.y1;
    return o.valid;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------------
void orc_median_stack_u16(const uint16_t* const* inputs, int numInputs, uint16_t* out,
                          long long numScalars, float sigmaThreshold) {
    int count = std::min(numInputs, 32);  // MedianStack_Scalar.cpp:12
    uint16_t v[32];
    for (long long i = 0; i < numScalars; ++i) {
        for (int k = 0; k < count; ++k) v[k] = inputs[k][i];
        out[i] = MedianOne(v, count, sigmaThreshold);
    }
}

// WarpPerspective_Scalar.cpp:8-45 + WarpPerspectiveKernels.h:16-43.
void orc_warp_perspective_u16(const uint16_t* src, int srcW, int srcH, uint16_t* dst, int dstW, int dstH,
                              const float* m, int offX, int offY) {
    for (int y = 0; y < dstH; ++y) {
        float Y = static_cast<float>(y + offY);
        for (int x = 0; x < dstW; ++x) {
            float X = static_cast<float>(x + offX);
            float u = (m[0] * X + m[1] * Y) + m[2];
            float v = (m[3] * X + m[4] * Y) + m[5];
            float q = (m[6] * X + m[7] * Y) + m[8];
            uint16_t* o = dst + (static_cast<size_t>(y) * dstW + x) * 3;
            if (std::fabs(q) < 1e-10f) { o[0] = o[1] = o[2] = 0; continue; }
            u = u / q;
            v = v / q;
            float fu = std::floor(u), fv = std::floor(v);
            // valid iff sx0 >= 0, sy0 >= 0, sx0+1 < srcW, sy0+1 < srcH; tested
            // on the floored float so huge / NaN coordinates are invalid
            // without an out-of-range float->int conversion.
            bool valid = fu >= 0.0f && fv >= 0.0f && fu <= static_cast<float>(srcW - 2) &&
                         fv <= static_cast<float>(srcH - 2);
            if (!valid) { o[0] = o[1] = o[2] = 0; continue; }
            int sx0 = static_cast<int>(fu), sy0 = static_cast<int>(fv);
            float fx = u - static_cast<float>(sx0);
            float fy = v - static_cast<float>(sy0);
            float omx = 1.0f - fx, omy = 1.0f - fy;
            float w00 = omx * omy, w10 = fx * omy, w01 = omx * fy, w11 = fx * fy;
            const uint16_t* r0 = src + (static_cast<size_t>(sy0) * srcW + sx0) * 3;
            const uint16_t* r1 = r0 + static_cast<size_t>(srcW) * 3;
            for (int c = 0; c < 3; ++c) {
                float a = w00 * static_cast<float>(r0[c]);
                float b = w10 * static_cast<float>(r0[3 + c]);
                float d = w01 * static_cast<float>(r1[c]);
                float e = w11 * static_cast<float>(r1[3 + c]);
                o[c] = SatU16(((a + b) + d) + e);
            }
        }
    }
}

// GainColorOps.cpp:16-26.
void orc_apply_gain_u16(uint16_t* data, long long n, float gain) {
    if (!data || n <= 0) return;
    if (std::fabs(gain - 1.0f) < 1e-6f) return;
    for (long long i = 0; i < n; ++i) data[i] = SatU16(static_cast<float>(data[i]) * gain);
}

void orc_demosaic_bayer_u16(const uint16_t* cfa, int w, int h, uint16_t black, const float* camMul,
                            const float* rgbCam12, uint32_t filters, uint16_t* out) {
    Demosaic<false>(cfa, w, h, black, camMul, rgbCam12, filters, out);
}

// Same pipeline; the colour-matrix result is clamped to [0,65535] and kept
// as float (the u16 entry point truncates exactly this value).
void orc_demosaic_bayer_f32(const uint16_t* cfa, int w, int h, uint16_t black, const float* camMul,
                            const float* rgbCam12, uint32_t filters, float* out) {
    Demosaic<true>(cfa, w, h, black, camMul, rgbCam12, filters, out);
}

// HomographyMath.cpp:94-116 (fp32 adjugate; |det| < 1e-12 fails).
int orc_invert_homography(const float* hm, float* inv) {
    float a = hm[0], b = hm[1], c = hm[2], d = hm[3], e = hm[4], f = hm[5], g = hm[6], i = hm[7], j = hm[8];
    float c0 = e * j - f * i, c1 = d * j - f * g, c2 = d * i - e * g;
    float det = (a * c0 - b * c1) + c * c2;
    if (std::fabs(det) < 1e-12f) return 0;
    float s = 1.0f / det;
    inv[0] = c0 * s;
    inv[1] = (c * i - b * j) * s;
    inv[2] = (b * f - c * e) * s;
    inv[3] = (f * g - d * j) * s;
    inv[4] = (a * j - c * g) * s;
    inv[5] = (c * d - a * f) * s;
    inv[6] = c2 * s;
    inv[7] = (b * g - a * i) * s;
    inv[8] = (a * e - b * d) * s;
    return 1;
}

// SeamBlend.cpp:206-242: two-pass 3x3 chamfer (1, 1.41421356f); seeds are
// uncovered pixels; kInf = w + h.
void orc_coverage_distance(const uint8_t* cov, int w, int h, float* d) {
    const float kD = 1.41421356f;
    const float kInf = static_cast<float>(w + h);
    size_t n = static_cast<size_t>(w) * h;
    for (size_t p = 0; p < n; ++p) d[p] = cov[p] ? kInf : 0.0f;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float b = d[Px(x, y, w)];
            if (x > 0) b = std::min(b, d[Px(x - 1, y, w)] + 1.0f);
            if (y > 0) b = std::min(b, d[Px(x, y - 1, w)] + 1.0f);
            if (x > 0 && y > 0) b = std::min(b, d[Px(x - 1, y - 1, w)] + kD);
            if (x < w - 1 && y > 0) b = std::min(b, d[Px(x + 1, y - 1, w)] + kD);
            d[Px(x, y, w)] = b;
        }
    for (int y = h - 1; y >= 0; --y)
        for (int x = w - 1; x >= 0; --x) {
            float b = d[Px(x, y, w)];
            if (x < w - 1) b = std::min(b, d[Px(x + 1, y, w)] + 1.0f);
            if (y < h - 1) b = std::min(b, d[Px(x, y + 1, w)] + 1.0f);
            if (x < w - 1 && y < h - 1) b = std::min(b, d[Px(x + 1, y + 1, w)] + kD);
            if (x > 0 && y < h - 1) b = std::min(b, d[Px(x - 1, y + 1, w)] + kD);
            d[Px(x, y, w)] = b;
        }
}

// SeamBlend.cpp:244-261: argmax distance > 0, ties -> lowest index.
void orc_ownership(const float* const* dist, int c, int w, int h, int32_t* owner) {
    size_t n = static_cast<size_t>(w) * h;
    for (size_t p = 0; p < n; ++p) {
        int best = -1;
        float bd = 0.0f;
        for (int i = 0; i < c; ++i)
            if (dist[i][p] > bd) { bd = dist[i][p]; best = i; }
        owner[p] = best;
    }
}

// SeamBlend.cpp:293-445: per-pair DP seam inside the overlap box.
// Boxes are passed as {x0,y0,x1,y1}; callers pass valid boxes only.
void orc_refine_seam_dp(const uint16_t* imgI, const uint16_t* imgJ, const float* distI, const float* distJ,
                        int w, int ci, int cj, const int* boxI, const int* boxJ, const int* ov,
                        int32_t* owner) {
    const int ox0 = ov[0], oy0 = ov[1], ox1 = ov[2], oy1 = ov[3];
    const int bw = ox1 - ox0 + 1, bh = oy1 - oy0 + 1;
    const bool rowScan = bw >= bh;
    float cI = rowScan ? static_cast<float>(boxI[0] + boxI[2]) * 0.5f : static_cast<float>(boxI[1] + boxI[3]) * 0.5f;
    float cJ = rowScan ? static_cast<float>(boxJ[0] + boxJ[2]) * 0.5f : static_cast<float>(boxJ[1] + boxJ[3]) * 0.5f;
    const bool iLow = cI <= cJ;
    const float kSmooth = 2.0f;
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
            if (distI[p] > 0.0f && distJ[p] > 0.0f) {
                if (hi[ln] < lo[ln]) { lo[ln] = t; hi[ln] = t; }
                else { lo[ln] = std::min(lo[ln], t); hi[ln] = std::max(hi[ln], t); }
            }
        }
    }
    std::vector<std::vector<float>> dp(lines);
    std::vector<std::vector<int>> bk(lines);
    std::vector<int> prevLine(lines, -1);
    int last = -1;
    for (int ln = 0; ln < lines; ++ln) {
        if (hi[ln] < lo[ln]) continue;
        int lw = hi[ln] - lo[ln] + 1;
        dp[ln].assign(lw, 0.0f);
        bk[ln].assign(lw, -1);
        for (int k = 0; k < lw; ++k) {
            int t = lo[ln] + k;
            size_t p = at(ln, t);
            float cost = 0.0f;
            for (int c = 0; c < 3; ++c)
                cost = cost + std::fabs(static_cast<float>(imgI[p * 3 + c]) - static_cast<float>(imgJ[p * 3 + c]));
            if (last < 0) { dp[ln][k] = cost; continue; }
            float best = std::numeric_limits<float>::max();
            int bp = -1;
            int plo = lo[last], pw = static_cast<int>(dp[last].size());
            for (int pk = 0; pk < pw; ++pk) {
                float cand = dp[last][pk] + kSmooth * std::fabs(static_cast<float>(t - (plo + pk)));
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
        float best = std::numeric_limits<float>::max();
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
            bool low = t < chosen[l];
            owner[p] = (low == iLow) ? ci : cj;
        }
    }
}

// SeamBlend.cpp:457-459: numBands limited by how often the chunk halves.
int orc_clamp_bands(int w, int h, int numBands) {
    int mb = 1;
    while (w > 1 && h > 1 && mb < numBands) { w = (w + 1) / 2; h = (h + 1) / 2; ++mb; }
    return std::min(numBands, std::max(1, mb));
}

// SeamBlend.cpp:447-546 given the final owner map and the coverage masks.
void orc_multiband_compose_u16(const uint16_t* const* img, const uint8_t* const* cov, const int32_t* owner,
                               int c, int w, int h, int numBands, uint16_t* out) {
    size_t n = static_cast<size_t>(w) * h;
    std::memset(out, 0, n * 3 * sizeof(uint16_t));
    int L = orc_clamp_bands(w, h, numBands);
    std::vector<bool> owns(c, false);
    for (size_t p = 0; p < n; ++p) if (owner[p] >= 0) owns[owner[p]] = true;
    std::vector<Level> band;
    bool init = false;
    for (int i = 0; i < c; ++i) {
        if (!owns[i]) continue;
        std::vector<float> wm(n, 0.0f), cm(n);
        for (size_t p = 0; p < n; ++p) { wm[p] = owner[p] == i ? 1.0f : 0.0f; cm[p] = cov[i][p] ? 1.0f : 0.0f; }
        std::vector<Level> wp, g, lap;
        GaussPyr(wm, w, h, 1, L, wp);
        HoleFilledPyr(img[i], cm, w, h, L, g);
        LaplacianPyr(g, lap);
        if (!init) {
            band.assign(L, Level{});
            for (int l = 0; l < L; ++l) { band[l].w = lap[l].w; band[l].h = lap[l].h; band[l].d.assign(lap[l].d.size(), 0.0f); }
            init = true;
        }
        for (int l = 0; l < L; ++l) {
            size_t m = static_cast<size_t>(band[l].w) * band[l].h;
            for (size_t p = 0; p < m; ++p) {
                float wt = wp[l].d[p];
                for (int k = 0; k < 3; ++k) band[l].d[p * 3 + k] = band[l].d[p * 3 + k] + wt * lap[l].d[p * 3 + k];
            }
        }
    }
    if (!init) return;
    std::vector<float> rec = band[L - 1].d;
    int rw = band[L - 1].w, rh = band[L - 1].h;
    for (int l = L - 2; l >= 0; --l) {
        std::vector<float> up;
        Expand(rec, rw, rh, 3, band[l].w, band[l].h, up);
        rec.resize(up.size());
        for (size_t p = 0; p < up.size(); ++p) rec[p] = band[l].d[p] + up[p];
        rw = band[l].w; rh = band[l].h;
    }
    for (size_t p = 0; p < n; ++p)
        for (int k = 0; k < 3; ++k) {
            float v = rec[p * 3 + k];
            float a = (v < 65535.0f) ? v : 65535.0f;   // (std::min)(65535.0f, v)
            float b = (0.0f < a) ? a : 0.0f;           // (std::max)(0.0f, a)
            out[p * 3 + k] = owner[p] >= 0 ? static_cast<uint16_t>(b + 0.5f) : 0;
        }
}

// SeamBlend.cpp:549-587 (BlendChunkContributors) with numBands exposed
// (the reference fixes it to kDefaultNumBands = 5).  ownerOut, when not
// null, receives the final owner map (after every pairwise seam pass).
void orc_blend_chunk_u16(const uint16_t* const* img, int c, int w, int h, int numBands, uint16_t* out,
                         int32_t* ownerOut) {
    size_t n = static_cast<size_t>(w) * h;
    std::memset(out, 0, n * 3 * sizeof(uint16_t));
    if (c <= 0) return;
    std::vector<std::vector<uint8_t>> cov(c, std::vector<uint8_t>(n));
    std::vector<std::vector<float>> dist(c, std::vector<float>(n));
    std::vector<Box> box(c);
    std::vector<const float*> dptr(c);
    std::vector<const uint8_t*> cptr(c);
    for (int i = 0; i < c; ++i) {
        for (size_t p = 0; p < n; ++p)
            cov[i][p] = (img[i][p * 3] != 0 || img[i][p * 3 + 1] != 0 || img[i][p * 3 + 2] != 0) ? 1 : 0;
        orc_coverage_distance(cov[i].data(), w, h, dist[i].data());
        box[i] = CoverBox(cov[i].data(), w, h);
        dptr[i] = dist[i].data();
        cptr[i] = cov[i].data();
    }
    std::vector<int32_t> owner(n);
    orc_ownership(dptr.data(), c, w, h, owner.data());
    for (int i = 0; i < c; ++i) {
        if (!box[i].valid) continue;
        for (int j = i + 1; j < c; ++j) {
            if (!box[j].valid) continue;
            Box o;
            if (!Intersect(box[i], box[j], o)) continue;
            int bi[4] = {box[i].x0, box[i].y0, box[i].x1, box[i].y1};
            int bj[4] = {box[j].x0, box[j].y0, box[j].x1, box[j].y1};
            int bo[4] = {o.x0, o.y0, o.x1, o.y1};
            orc_refine_seam_dp(img[i], img[j], dist[i].data(), dist[j].data(), w, i, j, bi, bj, bo, owner.data());
        }
    }
    if (ownerOut) std::memcpy(ownerOut, owner.data(), n * sizeof(int32_t));
    orc_multiband_compose_u16(img, cptr.data(), owner.data(), c, w, h, numBands, out);
}

// Coverage box helper exposed for tests: returns 1 and {x0,y0,x1,y1} when
// any pixel is covered.
int orc_cover_box(const uint8_t* cov, int w, int h, int* box) {
    Box b = CoverBox(cov, w, h);
    box[0] = b.x0; box[1] = b.y0; box[2] = b.x1; box[3] = b.y1;
    return b.valid ? 1 : 0;
}

}  // extern "C"
