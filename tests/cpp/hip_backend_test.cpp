// Drives HipComputeBackend (the C++ IComputeBackend implementation) the way
// the reference's callers do, through its C++ interface only.
//
//   hip_backend_test cpu          contract checks that need no GPU
//   hip_backend_test gpu OUTDIR   runs every GPU operation on seeded inputs
//                                 and writes inputs and outputs to OUTDIR
//                                 (raw little-endian arrays listed in
//                                 OUTDIR/manifest.txt) for the oracle
//                                 comparison in tests/test_cpp_backend.py
//
// Prints one "ok <check>" line per passed check; exits non-zero on the
// first failed one.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "HipBackendFactory.h"
#include "HipComputeBackend.h"

using namespace WindowsApp::Compute;
using WindowsApp::Hip::BackendSelection;
using WindowsApp::Hip::ComputeBackendKind;

namespace
{
    int g_fail = 0;

    void Check(bool cond, const char* what)
    {
        if (cond)
        {
            std::printf("ok %s\n", what);
            std::fflush(stdout);
            return;
        }
        std::printf("FAILED %s\n", what);
        std::fflush(stdout);
        g_fail = 1;
        std::exit(1);
    }

    // xorshift64* stream for the seeded inputs
    struct Rng
    {
        uint64_t s;
        explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 1) {}
        uint64_t Next()
        {
            s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
            return s * 0x2545F4914F6CDD1Dull;
        }
        uint16_t U16(uint32_t hi) { return static_cast<uint16_t>(Next() % (hi + 1)); }
    };

    // A host backend standing in for Core's CpuComputeBackend: it answers
    // the align operations (and, were they delegated, the optimize ones)
    // with fixed values and counts the calls.
    class FakeHost : public IComputeBackend
    {
    public:
        std::atomic<int> calls{0};
        bool inited = false;
        ComputeResult Initialize() override { inited = true; return ComputeResult::SUCCESS; }
        void Shutdown() override {}
        bool IsInitialized() const override { return inited; }
        GpuInfo GetGpuInfo() const override { GpuInfo g; std::snprintf(g.name, sizeof(g.name), "FakeCPU"); return g; }
        const char* GetLastError() const override { return "fake host error"; }
        ComputeResult WarpPerspective(const unsigned short*, int, int, unsigned short*, int, int, const float*, int, int) override { return ComputeResult::SUCCESS; }
        ComputeResult MedianStack(const unsigned short**, int, unsigned short*, int, int, float) override { return ComputeResult::SUCCESS; }
        ComputeResult ApplyGain(unsigned short*, int, float) override { return ComputeResult::SUCCESS; }
        ComputeResult DemosaicBayer(const unsigned short*, int, int, unsigned short, const float[4], const float[3][4], uint32_t, unsigned short*) override { return ComputeResult::SUCCESS; }
        // like CpuComputeBackend, the align operations answer CUDA_ERROR
        // until Initialize() has run
        ComputeResult DetectAndDescribeFeatures(const unsigned char*, int, int, FeaturePoint*, BriefDescriptor*, int* outCount, int) override
        { ++calls; if (!inited) return ComputeResult::CUDA_ERROR; *outCount = 7; return ComputeResult::SUCCESS; }
        ComputeResult MatchFeatures(const BriefDescriptor*, int, const BriefDescriptor*, int, MatchResult*, int* n, int, float) override
        { ++calls; if (!inited) return ComputeResult::CUDA_ERROR; *n = 3; return ComputeResult::INVALID_PARAM; }
        ComputeResult ComputeLabStats(const unsigned short*, int, int, double m[3], double s[3]) override
        { ++calls; for (int i = 0; i < 3; ++i) { m[i] = i + 0.5; s[i] = 2.0; } return ComputeResult::SUCCESS; }
        ComputeResult ApplyReinhardColorTransfer(unsigned short*, int, int, const double[3], const double[3], const double[3], const double[3]) override
        { ++calls; return ComputeResult::INVALID_PARAM; }
        ComputeResult BlockMatchAlign(const unsigned short*, const unsigned short*, int, int, int, int, TileOffset*, int, int) override { return ComputeResult::SUCCESS; }
        ComputeResult RobustMergeAccumulate(const unsigned short* const*, int, const TileOffset* const*, int, int, int, int, int, float, unsigned short*) override { return ComputeResult::SUCCESS; }
        ComputeResult TileFftMerge(const unsigned short* const*, int, const TileOffset* const*, int, int, int, int, int, float, unsigned short*) override { return ComputeResult::SUCCESS; }
        ComputeResult StructureTensorKernelRegression(const unsigned short* const*, int, const TileOffsetF* const*, int, int, int, int, int, int, float, unsigned short*) override { return ComputeResult::SUCCESS; }
    };

    void BurstOpsUnsupported(IComputeBackend& b)
    {
        unsigned short px[12] = {};
        TileOffset off[1];
        const unsigned short* frames[1] = {px};
        const TileOffset* offs[1] = {off};
        const TileOffsetF* offsF[1] = {nullptr};
        Check(b.BlockMatchAlign(px, px, 2, 2, 2, 1, off, 1, 1) == ComputeResult::NOT_SUPPORTED, "BlockMatchAlign NOT_SUPPORTED");
        Check(b.RobustMergeAccumulate(frames, 1, offs, 2, 2, 2, 1, 1, 1.0f, px) == ComputeResult::NOT_SUPPORTED, "RobustMergeAccumulate NOT_SUPPORTED");
        Check(b.TileFftMerge(frames, 1, offs, 2, 2, 2, 1, 1, 1.0f, px) == ComputeResult::NOT_SUPPORTED, "TileFftMerge NOT_SUPPORTED");
        Check(b.StructureTensorKernelRegression(frames, 1, offsF, 2, 2, 2, 1, 1, 2, 1.0f, px) == ComputeResult::NOT_SUPPORTED,
              "StructureTensorKernelRegression NOT_SUPPORTED");
    }

    int RunCpu()
    {
        HipComputeBackend b(0);
        Check(!b.IsInitialized(), "not initialised before Initialize");
        unsigned short px[12] = {};
        const unsigned short* in[1] = {px};
        const float h[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        const float cm[4] = {1, 1, 1, 1};
        const float rc[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
        Check(b.WarpPerspective(px, 2, 2, px, 2, 2, h, 0, 0) == ComputeResult::CUDA_ERROR &&
                  std::strcmp(b.GetLastError(), "Not initialized.") == 0, "WarpPerspective before Initialize");
        Check(b.MedianStack(in, 1, px, 2, 2) == ComputeResult::CUDA_ERROR, "MedianStack before Initialize");
        Check(b.ApplyGain(px, 12, 2.0f) == ComputeResult::CUDA_ERROR, "ApplyGain before Initialize");
        Check(b.DemosaicBayer(px, 2, 2, 0, cm, rc, 0x94949494u, px) == ComputeResult::CUDA_ERROR, "DemosaicBayer before Initialize");
        BurstOpsUnsupported(b);
        const ComputeResult r = b.Initialize();
        Check(r == ComputeResult::NO_GPU && std::strlen(b.GetLastError()) > 0 && !b.IsInitialized(),
              "Initialize without a gfx950 device reports NO_GPU");
        Check(WindowsApp::Hip::RecommendedChunkSize(288ull << 30) == 4096 &&
                  WindowsApp::Hip::RecommendedChunkSize(8ull << 30) == 2048 &&
                  WindowsApp::Hip::RecommendedChunkSize((8ull << 30) - 1) == 1024, "RecommendedChunkSize thresholds");
        {
            // the host delegate is initialised with the HIP backend, so its
            // align operations answer (the reference's factory initialises
            // every backend it returns)
            auto host = std::make_shared<FakeHost>();
            HipComputeBackend hb(0, host);
            (void)hb.Initialize();
            Check(host->inited, "Initialize initialises the host delegate");
        }
        auto fake = std::make_shared<FakeHost>();
        BackendSelection sel = WindowsApp::Hip::SelectComputeBackend(ComputeBackendKind::Auto, [&] { return fake; });
        Check(sel.kind == ComputeBackendKind::Cpu && !sel.usedGpu && sel.backend == fake && fake->inited &&
                  sel.statusMessage.find(L"No compatible GPU") == 0, "Auto falls back to the supplied CPU backend");
        return 0;
    }

    // ---------------------------------------------------------------- GPU
    struct Out
    {
        std::string dir;
        FILE* manifest = nullptr;
        template <typename T>
        void Put(const char* name, const char* dtype, const std::vector<T>& v, std::vector<int> shape)
        {
            const std::string path = dir + "/" + name + ".bin";
            FILE* f = std::fopen(path.c_str(), "wb");
            if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) Check(false, "write output file");
            std::fclose(f);
            std::fprintf(manifest, "%s %s", name, dtype);
            for (int d : shape) std::fprintf(manifest, " %d", d);
            std::fprintf(manifest, "\n");
        }
    };

    std::vector<unsigned short> Noise(Rng& r, size_t n, uint32_t hi)
    {
        std::vector<unsigned short> v(n);
        for (auto& x : v) x = r.U16(hi);
        return v;
    }

    int RunGpu(const char* dir)
    {
        Out out;
        out.dir = dir;
        out.manifest = std::fopen((std::string(dir) + "/manifest.txt").c_str(), "w");
        Check(out.manifest != nullptr, "open manifest");
        auto host = std::make_shared<FakeHost>();
        BackendSelection sel = WindowsApp::Hip::SelectComputeBackend(ComputeBackendKind::Hip, [&] { return host; });
        Check(sel.kind == ComputeBackendKind::Hip && sel.usedGpu && sel.recommendedMaxInFlight == 2 &&
                  sel.recommendedChunkSize == 4096, "factory selects the HIP backend");
        auto* b = dynamic_cast<HipComputeBackend*>(sel.backend.get());
        Check(b && b->IsInitialized(), "HipComputeBackend initialised");
        const GpuInfo gi = b->GetGpuInfo();
        Check(gi.totalMemory > (16ull << 30) && gi.multiProcessorCount > 0 && std::strstr(gi.name, "gfx950"),
              "GetGpuInfo describes the gfx950 device");
        std::printf("device: %s, %d CUs, %.1f GiB\n", gi.name, gi.multiProcessorCount, gi.totalMemory / double(1ull << 30));
        Rng rng(20261203);

        // validation order and messages (CpuComputeBackend.cpp:117-178)
        {
            unsigned short px[12] = {};
            const unsigned short* in33[33];
            for (auto& p : in33) p = px;
            Check(b->WarpPerspective(nullptr, 2, 2, px, 2, 2, nullptr, 0, 0) == ComputeResult::INVALID_PARAM &&
                      std::strcmp(b->GetLastError(), "Null pointer argument.") == 0, "warp null pointer");
            const float h[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            Check(b->WarpPerspective(px, 0, 2, px, 2, 2, h, 0, 0) == ComputeResult::INVALID_PARAM &&
                      std::strcmp(b->GetLastError(), "Invalid dimensions.") == 0, "warp bad dimensions");
            Check(b->MedianStack(in33, 33, px, 2, 2) == ComputeResult::INVALID_PARAM &&
                      std::strcmp(b->GetLastError(), "numInputs must be 1-32.") == 0, "median 33 inputs");
            Check(b->ApplyGain(px, 0, 2.0f) == ComputeResult::INVALID_PARAM &&
                      std::strcmp(b->GetLastError(), "Invalid pixel count.") == 0, "gain zero count");
            BurstOpsUnsupported(*b);
        }
        // align operations go to the host delegate
        {
            unsigned char img[12] = {};
            FeaturePoint pts[4];
            BriefDescriptor desc[4];
            MatchResult mr[4];
            int cnt = 0, nm = 0;
            Check(b->DetectAndDescribeFeatures(img, 2, 2, pts, desc, &cnt, 4) == ComputeResult::SUCCESS && cnt == 7 &&
                      host->calls == 1, "DetectAndDescribeFeatures delegated");
            Check(b->MatchFeatures(desc, 1, desc, 1, mr, &nm, 4) == ComputeResult::INVALID_PARAM &&
                      std::strcmp(b->GetLastError(), "fake host error") == 0, "delegate errors propagate");
            HipComputeBackend lone(0);
            Check(lone.Initialize() == ComputeResult::SUCCESS, "second instance initialises");
            Check(lone.DetectAndDescribeFeatures(img, 2, 2, pts, desc, &cnt, 4) == ComputeResult::NOT_SUPPORTED,
                  "no delegate: NOT_SUPPORTED");
            // the optimize colour operations run on the GPU, with the CPU backend's validation
            double m[3] = {50, 0, 0}, sd[3] = {10, 10, 10};
            unsigned short px[12] = {};
            Check(lone.ComputeLabStats(nullptr, 2, 2, m, sd) == ComputeResult::INVALID_PARAM &&
                      std::strcmp(lone.GetLastError(), "Null pointer argument.") == 0, "lab stats null");
            Check(b->ApplyReinhardColorTransfer(px, 2, 0, m, sd, m, sd) == ComputeResult::INVALID_PARAM &&
                      std::strcmp(b->GetLastError(), "Invalid dimensions.") == 0, "reinhard bad dimensions");
            Check(lone.ComputeLabStats(px, 2, 2, m, sd) == ComputeResult::SUCCESS && m[0] == 0.0 && sd[0] == 0.0,
                  "black image: L mean 0, std 0 without a delegate");
            Check(host->calls == 2, "optimize operations are not delegated");
        }
        // WarpPerspective
        const int sw = 301, sh = 203, dw = 257, dh = 190;
        const std::vector<unsigned short> src = Noise(rng, size_t(sw) * sh * 3, 65535);
        const float hinv[9] = {0.9962f, -0.0872f, 12.5f, 0.0872f, 0.9962f, -7.25f, 2e-5f, -1e-5f, 1.0f};
        std::vector<unsigned short> warp(size_t(dw) * dh * 3, 0);
        Check(b->WarpPerspective(src.data(), sw, sh, warp.data(), dw, dh, hinv, 5, -3) == ComputeResult::SUCCESS, "WarpPerspective runs");
        out.Put("warp_src", "u16", src, {sh, sw, 3});
        out.Put("warp_h", "f32", std::vector<float>(hinv, hinv + 9), {9});
        out.Put("warp_out", "u16", warp, {dh, dw, 3});
        // MedianStack (7 planes, some outliers)
        const int mw = 97, mh = 61, mn = 7;
        std::vector<std::vector<unsigned short>> planes;
        std::vector<const unsigned short*> pp;
        const std::vector<unsigned short> base = Noise(rng, size_t(mw) * mh * 3, 16383);
        for (int k = 0; k < mn; ++k)
        {
            std::vector<unsigned short> p = base;
            for (auto& x : p)
            {
                x = static_cast<unsigned short>(std::min<uint32_t>(65535, x + rng.U16(400)));
                if (rng.Next() % 50 == 0) x = rng.U16(65535);
            }
            planes.push_back(p);
        }
        for (auto& p : planes) pp.push_back(p.data());
        std::vector<unsigned short> med(size_t(mw) * mh * 3);
        Check(b->MedianStack(pp.data(), mn, med.data(), mw, mh, 2.0f) == ComputeResult::SUCCESS, "MedianStack runs");
        std::vector<unsigned short> allp;
        for (auto& p : planes) allp.insert(allp.end(), p.begin(), p.end());
        out.Put("median_in", "u16", allp, {mn, mh, mw, 3});
        out.Put("median_out", "u16", med, {mh, mw, 3});
        // ApplyGain, in place
        std::vector<unsigned short> g = Noise(rng, 3001, 65535);
        out.Put("gain_in", "u16", g, {3001});
        Check(b->ApplyGain(g.data(), static_cast<int>(g.size()), 1.37f) == ComputeResult::SUCCESS, "ApplyGain runs");
        out.Put("gain_out", "u16", g, {3001});
        // DemosaicBayer (RGGB)
        const int bw = 130, bh = 94;
        const std::vector<unsigned short> cfa = Noise(rng, size_t(bw) * bh, 16383);
        const float camMul[4] = {2.0f, 1.0f, 1.5f, 1.0f};
        const float rgbCam[3][4] = {{1.6f, -0.4f, -0.2f, 0.0f}, {-0.2f, 1.5f, -0.3f, 0.0f}, {0.05f, -0.5f, 1.45f, 0.0f}};
        std::vector<unsigned short> rgb(size_t(bw) * bh * 3);
        Check(b->DemosaicBayer(cfa.data(), bw, bh, 512, camMul, rgbCam, 0x94949494u, rgb.data()) == ComputeResult::SUCCESS,
              "DemosaicBayer runs");
        out.Put("demosaic_in", "u16", cfa, {bh, bw});
        out.Put("demosaic_out", "u16", rgb, {bh, bw, 3});
        // ComputeLabStats + ApplyReinhardColorTransfer (the optimize stage's
        // colour transfer, OptimizeExecutor.cpp:308-327)
        {
            const int lw = 211, lh = 97;
            const std::vector<unsigned short> tgt = Noise(rng, size_t(lw) * lh * 3, 65535);
            std::vector<unsigned short> ref = Noise(rng, size_t(lw) * lh * 3, 40000);
            double rm[3], rs[3], tm[3], ts[3];
            Check(b->ComputeLabStats(ref.data(), lw, lh, rm, rs) == ComputeResult::SUCCESS, "ComputeLabStats ref");
            Check(b->ComputeLabStats(tgt.data(), lw, lh, tm, ts) == ComputeResult::SUCCESS, "ComputeLabStats target");
            std::vector<unsigned short> moved = tgt;
            Check(b->ApplyReinhardColorTransfer(moved.data(), lw, lh, tm, ts, rm, rs) == ComputeResult::SUCCESS,
                  "ApplyReinhardColorTransfer runs");
            out.Put("lab_ref", "u16", ref, {lh, lw, 3});
            out.Put("lab_tgt", "u16", tgt, {lh, lw, 3});
            std::vector<double> stats(rm, rm + 3);
            stats.insert(stats.end(), rs, rs + 3);
            stats.insert(stats.end(), tm, tm + 3);
            stats.insert(stats.end(), ts, ts + 3);
            out.Put("lab_stats", "f64", stats, {4, 3});
            out.Put("reinhard_out", "u16", moved, {lh, lw, 3});
        }
        // BlendChunkContributors via the extension interface, as a render
        // driver finds it
        {
            const int cw = 200, ch = 120;
            std::vector<std::vector<unsigned short>> warped;
            const float shifts[3][2] = {{0.0f, 0.0f}, {-70.0f, 4.0f}, {-135.0f, -6.0f}};
            for (int k = 0; k < 3; ++k)
            {
                std::vector<unsigned short> s(size_t(120) * 110 * 3);
                for (size_t i = 0; i < s.size(); ++i) s[i] = static_cast<unsigned short>(8000 + (rng.Next() % 30000));
                const float hk[9] = {1, 0, shifts[k][0], 0, 1, shifts[k][1], 0, 0, 1};
                std::vector<unsigned short> wv(size_t(cw) * ch * 3, 0);
                Check(b->WarpPerspective(s.data(), 120, 110, wv.data(), cw, ch, hk, 0, 0) == ComputeResult::SUCCESS, "contributor warp");
                warped.push_back(wv);
            }
            std::vector<const unsigned short*> wp;
            for (auto& wv : warped) wp.push_back(wv.data());
            auto* blend = dynamic_cast<IChunkBlendBackend*>(sel.backend.get());
            Check(blend != nullptr, "IChunkBlendBackend discovered with dynamic_cast");
            std::vector<unsigned short> res;
            Check(blend->BlendChunkContributors(wp, cw, ch, res) == ComputeResult::SUCCESS && res.size() == size_t(cw) * ch * 3,
                  "BlendChunkContributors runs");
            std::vector<unsigned short> allw;
            for (auto& wv : warped) allw.insert(allw.end(), wv.begin(), wv.end());
            out.Put("blend_in", "u16", allw, {3, ch, cw, 3});
            out.Put("blend_out", "u16", res, {ch, cw, 3});
            std::vector<unsigned short> empty;
            Check(blend->BlendChunkContributors({}, 4, 3, empty) == ComputeResult::SUCCESS && empty.size() == 36 &&
                      empty[0] == 0, "no contributors: black chunk");
        }
        // one instance, called from four threads at once (TaskScheduler's
        // in-flight window): every result equals the sequential one
        {
            const int nt = 4, reps = 6;
            std::vector<std::vector<unsigned short>> seq(nt), par(nt);
            std::vector<std::array<float, 9>> hs(nt);
            for (int t = 0; t < nt; ++t)
            {
                hs[t] = {1.0f, 0.01f * t, 3.0f * t, -0.01f * t, 1.0f, -2.0f * t, 0.0f, 0.0f, 1.0f};
                seq[t].assign(size_t(dw) * dh * 3, 0);
                Check(b->WarpPerspective(src.data(), sw, sh, seq[t].data(), dw, dh, hs[t].data(), 0, 0) == ComputeResult::SUCCESS,
                      "sequential warp");
            }
            std::atomic<int> bad{0};
            std::vector<std::thread> th;
            for (int t = 0; t < nt; ++t)
                th.emplace_back([&, t] {
                    for (int r = 0; r < reps; ++r)
                    {
                        std::vector<unsigned short> o(size_t(dw) * dh * 3, 0);
                        if (b->WarpPerspective(src.data(), sw, sh, o.data(), dw, dh, hs[t].data(), 0, 0) != ComputeResult::SUCCESS ||
                            o != seq[t])
                            ++bad;
                    }
                });
            for (auto& x : th) x.join();
            Check(bad == 0, "concurrent calls on one instance match sequential results");
        }
        std::fclose(out.manifest);
        b->Shutdown();
        Check(!b->IsInitialized() && b->WarpPerspective(src.data(), sw, sh, warp.data(), dw, dh, hinv, 0, 0) == ComputeResult::CUDA_ERROR,
              "Shutdown returns to the uninitialised state");
        return 0;
    }
}

int main(int argc, char** argv)
{
    if (argc >= 2 && std::strcmp(argv[1], "cpu") == 0) return RunCpu();
    if (argc >= 3 && std::strcmp(argv[1], "gpu") == 0) return RunGpu(argv[2]);
    std::fprintf(stderr, "usage: %s cpu | gpu OUTDIR\n", argv[0]);
    return 2;
}
