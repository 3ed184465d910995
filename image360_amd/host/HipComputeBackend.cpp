// HipComputeBackend: IComputeBackend over the C ABI of libi360hip.so.
#include "HipComputeBackend.h"

#include <cstdio>
#include <cstring>

namespace WindowsApp { namespace Compute
{
    namespace
    {
        ComputeResult ToResult(int rc)
        {
            switch (rc)
            {
            case I360HIP_SUCCESS: return ComputeResult::SUCCESS;
            case I360HIP_NO_GPU: return ComputeResult::NO_GPU;
            case I360HIP_INVALID_PARAM: return ComputeResult::INVALID_PARAM;
            case I360HIP_OUT_OF_MEMORY: return ComputeResult::OUT_OF_MEMORY;
            case I360HIP_KERNEL_LAUNCH_FAILED: return ComputeResult::KERNEL_LAUNCH_FAILED;
            case I360HIP_NOT_SUPPORTED: return ComputeResult::NOT_SUPPORTED;
            default: return ComputeResult::CUDA_ERROR;
            }
        }
    }

    HipComputeBackend::HipComputeBackend(int device, std::shared_ptr<IComputeBackend> hostDelegate)
        : m_device(device), m_host(std::move(hostDelegate))
    {
    }

    HipComputeBackend::~HipComputeBackend()
    {
        Shutdown();
    }

    ComputeResult HipComputeBackend::Fail(ComputeResult code, const char* message)
    {
        std::lock_guard<std::mutex> lk(m_mu);
        std::snprintf(m_lastError, sizeof(m_lastError), "%s", message);
        return code;
    }

    ComputeResult HipComputeBackend::FromAbi(int rc)
    {
        if (rc == I360HIP_SUCCESS) return ComputeResult::SUCCESS;
        // the calling thread's own message: another thread's failure on the
        // same context cannot have replaced it since the call returned
        return Fail(ToResult(rc), i360hip_thread_error());
    }

    ComputeResult HipComputeBackend::Initialize()
    {
        if (m_ctx) return ComputeResult::SUCCESS;
        // The align operations go to the host delegate, which answers them
        // only once initialised (CpuComputeBackend returns CUDA_ERROR "Not
        // initialized." before); the reference's factory initialises every
        // backend it hands out (ComputeBackendFactory.cpp:108-136).
        if (m_host && !m_host->IsInitialized())
        {
            const ComputeResult hr = m_host->Initialize();
            if (hr != ComputeResult::SUCCESS) return Fail(hr, m_host->GetLastError());
        }
        i360hip_ctx* ctx = nullptr;
        const int rc = i360hip_create(m_device, &ctx);
        if (rc != I360HIP_SUCCESS) return Fail(ToResult(rc), i360hip_thread_error());
        i360hip_device_info di{};
        const int rc2 = i360hip_get_device_info(ctx, &di);
        if (rc2 != I360HIP_SUCCESS)
        {
            const ComputeResult r = Fail(ToResult(rc2), i360hip_thread_error());
            i360hip_destroy(ctx);
            return r;
        }
        GpuInfo info;
        info.deviceId = di.device_id;
        std::snprintf(info.name, sizeof(info.name), "%s (%s)", di.name, di.arch);
        info.totalMemory = di.total_memory;
        info.freeMemory = di.free_memory;
        info.computeMajor = di.compute_major;
        info.computeMinor = di.compute_minor;
        info.maxThreadsPerBlock = di.max_threads_per_block;
        info.multiProcessorCount = di.multiprocessor_count;
        info.hasTensorCores = di.has_matrix_cores != 0;
        m_info = info;
        m_ctx = ctx;
        return ComputeResult::SUCCESS;
    }

    void HipComputeBackend::Shutdown()
    {
        if (m_ctx)
        {
            i360hip_destroy(m_ctx);
            m_ctx = nullptr;
        }
    }

    bool HipComputeBackend::IsInitialized() const
    {
        return m_ctx != nullptr;
    }

    GpuInfo HipComputeBackend::GetGpuInfo() const
    {
        return m_info;
    }

    const char* HipComputeBackend::GetLastError() const
    {
        // A snapshot taken under the lock: another thread's failure may
        // replace m_lastError at any time, but not the text handed out
        // here (valid until this thread's next GetLastError call).
        thread_local char snapshot[sizeof(m_lastError)];
        std::lock_guard<std::mutex> lk(m_mu);
        std::memcpy(snapshot, m_lastError, sizeof(snapshot));
        return snapshot;
    }

    // ---- render / ingest operations (validation as CpuComputeBackend.cpp:112-193)

    ComputeResult HipComputeBackend::WarpPerspective(
        const unsigned short* srcData, int srcW, int srcH,
        unsigned short* dstData, int dstW, int dstH,
        const float* homography, int offsetX, int offsetY)
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (!srcData || !dstData || !homography) return Fail(ComputeResult::INVALID_PARAM, "Null pointer argument.");
        if (srcW <= 0 || srcH <= 0 || dstW <= 0 || dstH <= 0) return Fail(ComputeResult::INVALID_PARAM, "Invalid dimensions.");
        return FromAbi(i360hip_warp_perspective_u16(m_ctx, srcData, srcW, srcH, dstData, dstW, dstH, homography,
                                                    offsetX, offsetY));
    }

    ComputeResult HipComputeBackend::MedianStack(
        const unsigned short** inputs, int numInputs,
        unsigned short* output, int width, int height,
        float sigmaThreshold)
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (!inputs || !output) return Fail(ComputeResult::INVALID_PARAM, "Null pointer argument.");
        if (numInputs <= 0 || numInputs > I360HIP_MAX_STACK_INPUTS) return Fail(ComputeResult::INVALID_PARAM, "numInputs must be 1-32.");
        if (width <= 0 || height <= 0) return Fail(ComputeResult::INVALID_PARAM, "Invalid dimensions.");
        return FromAbi(i360hip_median_stack_u16(m_ctx, inputs, numInputs, output, width, height, sigmaThreshold));
    }

    ComputeResult HipComputeBackend::ApplyGain(unsigned short* data, int numPixels, float gain)
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (!data) return Fail(ComputeResult::INVALID_PARAM, "Null pointer argument.");
        if (numPixels <= 0) return Fail(ComputeResult::INVALID_PARAM, "Invalid pixel count.");
        return FromAbi(i360hip_apply_gain_u16(m_ctx, data, numPixels, gain));
    }

    ComputeResult HipComputeBackend::DemosaicBayer(
        const unsigned short* cfaData, int width, int height,
        unsigned short blackLevel, const float camMul[4], const float rgbCam[3][4],
        uint32_t filters, unsigned short* rgbOut)
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (!cfaData || !camMul || !rgbCam || !rgbOut) return Fail(ComputeResult::INVALID_PARAM, "Null pointer argument.");
        if (width <= 0 || height <= 0) return Fail(ComputeResult::INVALID_PARAM, "Invalid dimensions.");
        return FromAbi(i360hip_demosaic_bayer_u16(m_ctx, cfaData, width, height, blackLevel, camMul, &rgbCam[0][0],
                                                  filters, rgbOut));
    }

    // ---- align operations: served by the host delegate

    ComputeResult HipComputeBackend::Delegated(const char* op)
    {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "%s: no host backend was given to serve it on the HIP backend.", op);
        return Fail(ComputeResult::NOT_SUPPORTED, msg);
    }

    ComputeResult HipComputeBackend::DetectAndDescribeFeatures(
        const unsigned char* rgbImage, int width, int height,
        FeaturePoint* outPoints, BriefDescriptor* outDescriptors, int* outCount, int maxPoints)
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (!m_host) return Delegated("DetectAndDescribeFeatures");
        const ComputeResult r = m_host->DetectAndDescribeFeatures(rgbImage, width, height, outPoints, outDescriptors,
                                                                  outCount, maxPoints);
        if (r != ComputeResult::SUCCESS) Fail(r, m_host->GetLastError());
        return r;
    }

    ComputeResult HipComputeBackend::MatchFeatures(
        const BriefDescriptor* descA, int countA,
        const BriefDescriptor* descB, int countB,
        MatchResult* outMatches, int* outMatchCount, int maxMatches,
        float ratioThreshold)
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (!m_host) return Delegated("MatchFeatures");
        const ComputeResult r = m_host->MatchFeatures(descA, countA, descB, countB, outMatches, outMatchCount,
                                                      maxMatches, ratioThreshold);
        if (r != ComputeResult::SUCCESS) Fail(r, m_host->GetLastError());
        return r;
    }

    // ---- optimize-stage colour operations: on the GPU (GainColorOps.cpp:103-157)

    ComputeResult HipComputeBackend::ComputeLabStats(
        const unsigned short* rgb, int width, int height, double outMean[3], double outStd[3])
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (!rgb || !outMean || !outStd) return Fail(ComputeResult::INVALID_PARAM, "Null pointer argument.");
        if (width <= 0 || height <= 0) return Fail(ComputeResult::INVALID_PARAM, "Invalid dimensions.");
        return FromAbi(i360hip_compute_lab_stats(m_ctx, rgb, width, height, outMean, outStd));
    }

    ComputeResult HipComputeBackend::ApplyReinhardColorTransfer(
        unsigned short* rgbInOut, int width, int height,
        const double srcMean[3], const double srcStd[3],
        const double refMean[3], const double refStd[3])
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (!rgbInOut || !srcMean || !srcStd || !refMean || !refStd)
            return Fail(ComputeResult::INVALID_PARAM, "Null pointer argument.");
        if (width <= 0 || height <= 0) return Fail(ComputeResult::INVALID_PARAM, "Invalid dimensions.");
        return FromAbi(i360hip_apply_reinhard_u16(m_ctx, rgbInOut, width, height, srcMean, srcStd, refMean, refStd));
    }

    // ---- burst-mode operations: not on this backend

    ComputeResult HipComputeBackend::BlockMatchAlign(
        const unsigned short*, const unsigned short*, int, int, int, int, TileOffset*, int, int)
    {
        return Fail(ComputeResult::NOT_SUPPORTED, "BlockMatchAlign is not implemented on the HIP backend (CPU backend only).");
    }

    ComputeResult HipComputeBackend::RobustMergeAccumulate(
        const unsigned short* const*, int, const TileOffset* const*, int, int, int, int, int, float, unsigned short*)
    {
        return Fail(ComputeResult::NOT_SUPPORTED, "RobustMergeAccumulate is not implemented on the HIP backend (CPU backend only).");
    }

    ComputeResult HipComputeBackend::TileFftMerge(
        const unsigned short* const*, int, const TileOffset* const*, int, int, int, int, int, float, unsigned short*)
    {
        return Fail(ComputeResult::NOT_SUPPORTED, "TileFftMerge is not implemented on the HIP backend (CPU backend only).");
    }

    ComputeResult HipComputeBackend::StructureTensorKernelRegression(
        const unsigned short* const*, int, const TileOffsetF* const*, int, int, int, int, int, int, float,
        unsigned short*)
    {
        return Fail(ComputeResult::NOT_SUPPORTED,
                    "StructureTensorKernelRegression is not implemented on the HIP backend (CPU backend only).");
    }

    // ---- extensions

    ComputeResult HipComputeBackend::BlendChunkContributors(
        const std::vector<const unsigned short*>& warpedBuffers,
        int width, int height, std::vector<unsigned short>& outResult, int numBands)
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        if (width <= 0 || height <= 0) return Fail(ComputeResult::INVALID_PARAM, "Invalid dimensions.");
        for (const unsigned short* p : warpedBuffers)
            if (!p) return Fail(ComputeResult::INVALID_PARAM, "Null pointer argument.");
        outResult.assign(static_cast<size_t>(width) * height * 3, 0);
        if (warpedBuffers.empty()) return ComputeResult::SUCCESS;  // all zero, as SeamBlend.cpp:552-555
        return FromAbi(i360hip_blend_chunk_u16(m_ctx, warpedBuffers.data(), static_cast<int>(warpedBuffers.size()),
                                               width, height, numBands, outResult.data(), nullptr));
    }

    ComputeResult HipComputeBackend::RenderChunks(const i360hip_frame* frames, int numFrames,
                                                  const i360hip_chunk* chunks, int numChunks,
                                                  const int32_t* contributors, int numBands, int mode, void* stream)
    {
        if (!m_ctx) return Fail(ComputeResult::CUDA_ERROR, "Not initialized.");
        return FromAbi(i360hip_render_chunks(m_ctx, frames, numFrames, chunks, numChunks, contributors, numBands,
                                             mode, stream));
    }
}}
