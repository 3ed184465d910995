// HipComputeBackend: the MI355X (gfx950) implementation of
// WindowsApp::Compute::IComputeBackend (WindowsApp.Compute/HeaderFiles/
// IComputeBackend.h:16-169), a drop-in beside CudaPipeline, VulkanPipeline
// and CpuComputeBackend.  Plain C++17 host code over the C ABI of
// libi360hip.so (include/i360hip.h); it includes no HIP or torch header.
//
// Contract, as the reference's backends keep it:
//  * every pointer is caller-owned host memory, outputs pre-allocated;
//  * calls are synchronous (results valid on return);
//  * results are ComputeResult codes, the text of the last failure is
//    returned by GetLastError() (one 512-byte record per instance, shared by
//    its callers as in every reference backend; the pointer returned is a
//    snapshot taken under the instance's lock, valid until the calling
//    thread's next GetLastError call);
//  * argument validation runs in CpuComputeBackend's order and with its
//    messages (CpuComputeBackend.cpp:112-193);
//  * one instance may be called from several threads at once (TaskScheduler
//    runs two GPU tasks in flight, TaskScheduler.cpp:67-71): the C ABI
//    context serialises its calls, this class guards its own state.
//
// Warp / median / gain / demosaic and the optimize stage's colour
// operations (Lab statistics, Reinhard transfer) run on the GPU.  The align
// operations (features, matching) are not on the render path: like
// VulkanPipeline (VulkanPipeline.h:18-23) this backend hands them to a host
// backend supplied at construction (in the reference build, Core's
// CpuComputeBackend) and reports NOT_SUPPORTED when none was given.  The burst-mode operations report NOT_SUPPORTED, as
// CudaPipeline does (CudaPipeline.cpp:794-839).
//
// Beyond IComputeBackend it implements IChunkBlendBackend, the GPU form of
// Kernels::SeamBlend::BlendChunkContributors (SeamBlendKernels.h:91-92), and
// exposes the GPU-resident float chunk render of the C ABI.
#pragma once

#include <memory>
#include <mutex>
#include <vector>

#include "IComputeBackend.h"
#include "i360hip.h"

namespace WindowsApp { namespace Compute
{
    // Optional extension of a backend: the whole per-chunk seam blend.  A
    // render driver discovers it with dynamic_cast and falls back to the
    // host Kernels::SeamBlend::BlendChunkContributors when absent.
    class IChunkBlendBackend
    {
    public:
        virtual ~IChunkBlendBackend() = default;
        // Same arguments and result as Kernels::SeamBlend::
        // BlendChunkContributors (SeamBlend.cpp:549-587): RGB48 warped
        // contributors of one chunk, (0,0,0) = no data; outResult is resized
        // to width*height*3 and holds the blended chunk.  numBands 5 is the
        // reference's fixed band count (SeamBlendKernels.h:10).
        virtual ComputeResult BlendChunkContributors(
            const std::vector<const unsigned short*>& warpedBuffers,
            int width, int height, std::vector<unsigned short>& outResult,
            int numBands = 5) = 0;
    };

    class HipComputeBackend : public IComputeBackend, public IChunkBlendBackend
    {
    public:
        // device: HIP device index.  hostDelegate: backend that serves the
        // align operations (may be null).
        explicit HipComputeBackend(int device = 0, std::shared_ptr<IComputeBackend> hostDelegate = nullptr);
        ~HipComputeBackend() override;

        HipComputeBackend(const HipComputeBackend&) = delete;
        HipComputeBackend& operator=(const HipComputeBackend&) = delete;

        ComputeResult Initialize() override;
        void Shutdown() override;
        bool IsInitialized() const override;
        GpuInfo GetGpuInfo() const override;
        const char* GetLastError() const override;

        ComputeResult WarpPerspective(
            const unsigned short* srcData, int srcW, int srcH,
            unsigned short* dstData, int dstW, int dstH,
            const float* homography, int offsetX, int offsetY) override;

        ComputeResult MedianStack(
            const unsigned short** inputs, int numInputs,
            unsigned short* output, int width, int height,
            float sigmaThreshold = 2.0f) override;

        ComputeResult ApplyGain(
            unsigned short* data, int numPixels, float gain) override;

        ComputeResult DemosaicBayer(
            const unsigned short* cfaData, int width, int height,
            unsigned short blackLevel, const float camMul[4], const float rgbCam[3][4],
            uint32_t filters, unsigned short* rgbOut) override;

        ComputeResult DetectAndDescribeFeatures(
            const unsigned char* rgbImage, int width, int height,
            FeaturePoint* outPoints, BriefDescriptor* outDescriptors, int* outCount, int maxPoints) override;

        ComputeResult MatchFeatures(
            const BriefDescriptor* descA, int countA,
            const BriefDescriptor* descB, int countB,
            MatchResult* outMatches, int* outMatchCount, int maxMatches,
            float ratioThreshold = 0.75f) override;

        ComputeResult ComputeLabStats(
            const unsigned short* rgb, int width, int height, double outMean[3], double outStd[3]) override;

        ComputeResult ApplyReinhardColorTransfer(
            unsigned short* rgbInOut, int width, int height,
            const double srcMean[3], const double srcStd[3],
            const double refMean[3], const double refStd[3]) override;

        ComputeResult BlockMatchAlign(
            const unsigned short* refData, const unsigned short* srcData,
            int width, int height, int tileSize, int searchRadius,
            TileOffset* outOffsets, int tilesX, int tilesY) override;

        ComputeResult RobustMergeAccumulate(
            const unsigned short* const* frames, int numFrames,
            const TileOffset* const* perFrameOffsets,
            int width, int height, int tileSize, int tilesX, int tilesY,
            float sigma, unsigned short* output) override;

        ComputeResult TileFftMerge(
            const unsigned short* const* frames, int numFrames,
            const TileOffset* const* perFrameOffsets,
            int width, int height, int tileSize, int tilesX, int tilesY,
            float noiseVariance, unsigned short* output) override;

        ComputeResult StructureTensorKernelRegression(
            const unsigned short* const* frames, int numFrames,
            const TileOffsetF* const* perFrameOffsets,
            int width, int height, int tileSize, int tilesX, int tilesY,
            int scaleFactor, float noiseVariance, unsigned short* output) override;

        ComputeResult BlendChunkContributors(
            const std::vector<const unsigned short*>& warpedBuffers,
            int width, int height, std::vector<unsigned short>& outResult,
            int numBands = 5) override;

        // GPU-resident float render of many chunks (i360hip_render_chunks):
        // device pointers, asynchronous on `stream` (null = the backend's
        // own stream).  mode: I360HIP_RENDER_*.
        ComputeResult RenderChunks(const i360hip_frame* frames, int numFrames,
                                   const i360hip_chunk* chunks, int numChunks,
                                   const int32_t* contributors, int numBands, int mode, void* stream);

        // The C ABI context (null before Initialize), for the device-pointer
        // entry points of i360hip.h.
        i360hip_ctx* Context() const { return m_ctx; }

    private:
        ComputeResult Fail(ComputeResult code, const char* message);
        ComputeResult FromAbi(int rc);
        ComputeResult Delegated(const char* op);

        int m_device;
        std::shared_ptr<IComputeBackend> m_host;
        i360hip_ctx* m_ctx = nullptr;
        GpuInfo m_info;
        mutable std::mutex m_mu;  // guards m_lastError
        char m_lastError[512] = {};
    };
}}
