// Concurrent calls on one HipComputeBackend, for sanitizer builds of the C++
// host layer on the CPU (linked against tests/cpp/abi_stub.cpp instead of
// the GPU library; built by tests/test_cpp_backend.py with
// -fsanitize=thread, then with -fsanitize=address,undefined).
//
// Eight threads share one instance (TaskScheduler calls one backend from
// its std::async workers, TaskScheduler.cpp:67-71) and mix succeeding
// calls, failing C ABI calls with thread-specific messages, argument
// validation failures and delegated align operations.  Checks: every
// result is correct, and every GetLastError text is one whole message that
// some call produced (never torn, never empty).  The sanitizer reports any
// data race or memory error on top.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "HipBackendFactory.h"
#include "HipComputeBackend.h"

using namespace WindowsApp::Compute;

namespace
{
    class Host : public IComputeBackend
    {
    public:
        std::atomic<bool> inited{false};
        ComputeResult Initialize() override { inited = true; return ComputeResult::SUCCESS; }
        void Shutdown() override {}
        bool IsInitialized() const override { return inited; }
        GpuInfo GetGpuInfo() const override { return GpuInfo{}; }
        const char* GetLastError() const override { return "host error"; }
        ComputeResult WarpPerspective(const unsigned short*, int, int, unsigned short*, int, int, const float*, int, int) override { return ComputeResult::SUCCESS; }
        ComputeResult MedianStack(const unsigned short**, int, unsigned short*, int, int, float) override { return ComputeResult::SUCCESS; }
        ComputeResult ApplyGain(unsigned short*, int, float) override { return ComputeResult::SUCCESS; }
        ComputeResult DemosaicBayer(const unsigned short*, int, int, unsigned short, const float[4], const float[3][4], uint32_t, unsigned short*) override { return ComputeResult::SUCCESS; }
        ComputeResult DetectAndDescribeFeatures(const unsigned char*, int, int, FeaturePoint*, BriefDescriptor*, int* n, int) override
        { if (!inited) return ComputeResult::CUDA_ERROR; *n = 5; return ComputeResult::SUCCESS; }
        ComputeResult MatchFeatures(const BriefDescriptor*, int, const BriefDescriptor*, int, MatchResult*, int* n, int, float) override
        { *n = 0; return ComputeResult::INVALID_PARAM; }
        ComputeResult ComputeLabStats(const unsigned short*, int, int, double[3], double[3]) override { return ComputeResult::SUCCESS; }
        ComputeResult ApplyReinhardColorTransfer(unsigned short*, int, int, const double[3], const double[3], const double[3], const double[3]) override { return ComputeResult::SUCCESS; }
        ComputeResult BlockMatchAlign(const unsigned short*, const unsigned short*, int, int, int, int, TileOffset*, int, int) override { return ComputeResult::SUCCESS; }
        ComputeResult RobustMergeAccumulate(const unsigned short* const*, int, const TileOffset* const*, int, int, int, int, int, float, unsigned short*) override { return ComputeResult::SUCCESS; }
        ComputeResult TileFftMerge(const unsigned short* const*, int, const TileOffset* const*, int, int, int, int, int, float, unsigned short*) override { return ComputeResult::SUCCESS; }
        ComputeResult StructureTensorKernelRegression(const unsigned short* const*, int, const TileOffsetF* const*, int, int, int, int, int, int, float, unsigned short*) override { return ComputeResult::SUCCESS; }
    };

    bool WellFormed(const char* e, int threads)
    {
        if (std::strcmp(e, "Null pointer argument.") == 0 || std::strcmp(e, "numInputs must be 1-32.") == 0 ||
            std::strcmp(e, "host error") == 0)
            return true;
        int off = -1;
        char tail = 0;
        if (std::sscanf(e, "stub: warp failed for offset %d%c", &off, &tail) == 1)
            return off >= 1000 && off < 1000 + threads;
        return false;
    }
}

int main()
{
    auto host = std::make_shared<Host>();
    HipComputeBackend b(0, host);
    if (b.Initialize() != ComputeResult::SUCCESS || !host->inited)
    {
        std::printf("FAILED initialise (%s)\n", b.GetLastError());
        return 1;
    }
    const int threads = 8, reps = 300;
    const int sw = 37, sh = 23, dw = 31, dh = 19;
    std::vector<unsigned short> src(size_t(sw) * sh * 3);
    for (size_t i = 0; i < src.size(); ++i) src[i] = static_cast<unsigned short>(i * 7919u);
    std::atomic<int> bad{0}, torn{0};
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t)
        th.emplace_back([&, t] {
            const float h[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            std::vector<unsigned short> out(size_t(dw) * dh * 3);
            for (int r = 0; r < reps; ++r)
            {
                switch ((r + t) % 5)
                {
                case 0:  // succeeds: the stub copies the top-left window
                {
                    if (b.WarpPerspective(src.data(), sw, sh, out.data(), dw, dh, h, 0, 0) != ComputeResult::SUCCESS ||
                        out[(size_t(5) * dw + 7) * 3 + 1] != src[(size_t(5) * sw + 7) * 3 + 1])
                        ++bad;
                    break;
                }
                case 1:  // the C ABI fails with this thread's message
                    if (b.WarpPerspective(src.data(), sw, sh, out.data(), dw, dh, h, 1000 + t, 0) !=
                        ComputeResult::KERNEL_LAUNCH_FAILED)
                        ++bad;
                    break;
                case 2:  // validation failure in the host layer
                    if (b.WarpPerspective(nullptr, sw, sh, out.data(), dw, dh, h, 0, 0) != ComputeResult::INVALID_PARAM)
                        ++bad;
                    break;
                case 3:
                {
                    const unsigned short* in[33];
                    for (auto& p : in) p = src.data();
                    if (b.MedianStack(in, 33, out.data(), 2, 2) != ComputeResult::INVALID_PARAM) ++bad;
                    break;
                }
                default:  // delegated align operation (the delegate was initialised)
                {
                    int n = 0;
                    if (b.DetectAndDescribeFeatures(nullptr, 1, 1, nullptr, nullptr, &n, 4) != ComputeResult::SUCCESS ||
                        n != 5)
                        ++bad;
                    MatchResult mr[1];
                    (void)b.MatchFeatures(nullptr, 0, nullptr, 0, mr, &n, 1);
                    break;
                }
                }
                const char* e = b.GetLastError();
                if (r > 4 && !WellFormed(e, threads)) ++torn;
            }
        });
    for (auto& x : th) x.join();
    std::printf("%s concurrent results (%d wrong)\n", bad == 0 ? "ok" : "FAILED", bad.load());
    std::printf("%s error texts whole (%d torn or foreign)\n", torn == 0 ? "ok" : "FAILED", torn.load());
    b.Shutdown();
    return (bad == 0 && torn == 0) ? 0 : 1;
}
