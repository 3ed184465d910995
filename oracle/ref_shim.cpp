// C entry points over the reference's own hot-path sources, compiled where
// they lie (see oracle/Makefile, target `ref`).  TEST INFRASTRUCTURE ONLY:
// used to pin the oracle restatement and as the timed CPU baseline.
// This file contains no reference code: it only calls the reference
// functions declared in the reference headers.

#include <cstdint>
#include <vector>

#include "HeaderFiles/BayerDemosaicKernels.h"
#include "HeaderFiles/CpuSimdDetect.h"
#include "HeaderFiles/GainColorOps.h"
#include "HeaderFiles/HomographyMath.h"
#include "HeaderFiles/MedianStackKernels.h"
#include "HeaderFiles/SeamBlendKernels.h"
#include "HeaderFiles/WarpPerspectiveKernels.h"

namespace K = WindowsApp::Core::Kernels;

extern "C" {

// tier: 0 scalar, 1 avx2, 2 avx512
int ref_simd_tier() {
    auto t = WindowsApp::Core::DetectCpuSimdTier();
    if (t == WindowsApp::Core::SimdTier::Avx512) return 2;
    if (t == WindowsApp::Core::SimdTier::Avx2) return 1;
    return 0;
}

void ref_median_stack(int tier, const uint16_t* const* in, int n, uint16_t* out, int w, int h, float sigma) {
    auto p = const_cast<const unsigned short**>(reinterpret_cast<const unsigned short* const*>(in));
    if (tier == 2) K::Avx512::MedianStack(p, n, out, w, h, sigma);
    else if (tier == 1) K::Avx2::MedianStack(p, n, out, w, h, sigma);
    else K::Scalar::MedianStack(p, n, out, w, h, sigma);
}

void ref_warp_perspective(int tier, const uint16_t* src, int sw, int sh, uint16_t* dst, int dw, int dh,
                          const float* invH, int ox, int oy) {
    if (tier == 2) K::Avx512::WarpPerspective(src, sw, sh, dst, dw, dh, invH, ox, oy);
    else if (tier == 1) K::Avx2::WarpPerspective(src, sw, sh, dst, dw, dh, invH, ox, oy);
    else K::Scalar::WarpPerspective(src, sw, sh, dst, dw, dh, invH, ox, oy);
}

void ref_demosaic_bayer(int tier, const uint16_t* cfa, int w, int h, uint16_t black, const float* camMul,
                        const float* rgbCam12, uint32_t filters, uint16_t* out) {
    const float(*cam)[4] = reinterpret_cast<const float(*)[4]>(rgbCam12);
    if (tier == 2) K::Avx512::DemosaicBayer(cfa, w, h, black, camMul, cam, filters, out);
    else if (tier == 1) K::Avx2::DemosaicBayer(cfa, w, h, black, camMul, cam, filters, out);
    else K::Scalar::DemosaicBayer(cfa, w, h, black, camMul, cam, filters, out);
}

void ref_apply_gain(uint16_t* data, int n, float gain) { WindowsApp::Core::ApplyGainCpu(data, n, gain); }

void ref_lab_stats(const uint16_t* rgb, int w, int h, double* mean, double* sd) {
    WindowsApp::Core::ComputeLabStatsCpu(rgb, w, h, mean, sd);
}

void ref_reinhard(uint16_t* rgb, int w, int h, const double* sm, const double* ss, const double* rm, const double* rs) {
    WindowsApp::Core::ApplyReinhardColorTransferCpu(rgb, w, h, sm, ss, rm, rs);
}

int ref_invert_homography(const float* h, float* inv) { return WindowsApp::Core::InvertHomography3x3(h, inv) ? 1 : 0; }

void ref_blend_chunk(const uint16_t* const* warped, int c, int w, int h, uint16_t* out) {
    std::vector<const unsigned short*> ptrs(warped, warped + c);
    std::vector<unsigned short> res;
    K::SeamBlend::BlendChunkContributors(ptrs, w, h, res);
    for (size_t i = 0; i < res.size(); ++i) out[i] = res[i];
}

void ref_coverage_distance(const uint8_t* cov, int w, int h, float* dist) {
    std::vector<unsigned char> c(cov, cov + static_cast<size_t>(w) * h);
    std::vector<float> d;
    K::SeamBlend::Detail::ComputeCoverageDistance(c, w, h, d);
    for (size_t i = 0; i < d.size(); ++i) dist[i] = d[i];
}

void ref_multiband_compose(const uint16_t* const* warped, const float* const* dist, const int32_t* owner, int c,
                           int w, int h, int numBands, uint16_t* out) {
    size_t n = static_cast<size_t>(w) * h;
    std::vector<const unsigned short*> ptrs(warped, warped + c);
    std::vector<std::vector<float>> d(c);
    for (int i = 0; i < c; ++i) d[i].assign(dist[i], dist[i] + n);
    std::vector<int> own(owner, owner + n);
    std::vector<unsigned short> res;
    K::SeamBlend::Detail::MultiBandCompose(ptrs, d, own, w, h, numBands, res);
    for (size_t i = 0; i < res.size(); ++i) out[i] = res[i];
}

// BlendChunkContributors with a chosen band count: the same Detail calls in
// the order of SeamBlend.cpp:549-587, ending in MultiBandCompose(numBands).
void ref_blend_owner(const uint16_t* const* warped, int c, int w, int h, int32_t* ownerOut);
void ref_blend_chunk_bands(const uint16_t* const* warped, int c, int w, int h, int numBands, uint16_t* out) {
    size_t n = static_cast<size_t>(w) * h;
    std::vector<int32_t> owner(n);
    ref_blend_owner(warped, c, w, h, owner.data());
    std::vector<const unsigned short*> ptrs(warped, warped + c);
    std::vector<std::vector<float>> dist(c);
    for (int i = 0; i < c; ++i) {
        std::vector<unsigned char> cov(n);
        for (size_t p = 0; p < n; ++p) cov[p] = (warped[i][p * 3] | warped[i][p * 3 + 1] | warped[i][p * 3 + 2]) ? 1 : 0;
        K::SeamBlend::Detail::ComputeCoverageDistance(cov, w, h, dist[i]);
    }
    std::vector<int> own(owner.begin(), owner.end());
    std::vector<unsigned short> res;
    K::SeamBlend::Detail::MultiBandCompose(ptrs, dist, own, w, h, numBands, res);
    for (size_t i = 0; i < res.size(); ++i) out[i] = res[i];
}

// Full BlendChunkContributors pipeline, but returning the owner map that
// feeds MultiBandCompose (reassembled from the public Detail functions in
// the same order as SeamBlend.cpp:549-587).
void ref_blend_owner(const uint16_t* const* warped, int c, int w, int h, int32_t* ownerOut) {
    namespace D = K::SeamBlend::Detail;
    size_t n = static_cast<size_t>(w) * h;
    std::vector<std::vector<unsigned char>> cov(c);
    std::vector<std::vector<float>> dist(c);
    std::vector<D::BBox> box(c);
    for (int i = 0; i < c; ++i) {
        cov[i].assign(n, 0);
        for (size_t p = 0; p < n; ++p)
            cov[i][p] = (warped[i][p * 3] | warped[i][p * 3 + 1] | warped[i][p * 3 + 2]) ? 1 : 0;
        D::ComputeCoverageDistance(cov[i], w, h, dist[i]);
        box[i] = D::ComputeCoverageBBox(cov[i], w, h);
    }
    std::vector<int> owner;
    D::ComputeOwnership(dist, w, h, owner);
    for (int i = 0; i < c; ++i) {
        if (!box[i].valid) continue;
        for (int j = i + 1; j < c; ++j) {
            if (!box[j].valid) continue;
            D::BBox o;
            if (!D::IntersectBBox(box[i], box[j], o)) continue;
            D::RefineSeamDP(warped[i], warped[j], dist[i], dist[j], w, h, i, j, box[i], box[j], o, owner);
        }
    }
    for (size_t p = 0; p < n; ++p) ownerOut[p] = owner[p];
}

}  // extern "C"
