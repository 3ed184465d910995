#include "HipBackendFactory.h"

#include <algorithm>
#include <thread>

#include "HipComputeBackend.h"

namespace WindowsApp { namespace Hip
{
    namespace
    {
        std::wstring Widen(const char* s)
        {
            std::wstring w;
            for (; s && *s; ++s) w.push_back(static_cast<wchar_t>(static_cast<unsigned char>(*s)));
            return w;
        }
    }

    int RecommendedChunkSize(uint64_t totalVramBytes)
    {
        constexpr uint64_t kGiB = 1024ull * 1024 * 1024;
        if (totalVramBytes >= 16 * kGiB) return 4096;
        if (totalVramBytes >= 8 * kGiB) return 2048;
        return 1024;
    }

    bool TryHip(BackendSelection& sel, std::wstring& lastError,
                std::shared_ptr<Compute::IComputeBackend> hostDelegate, int device)
    {
        auto hip = std::make_shared<Compute::HipComputeBackend>(device, std::move(hostDelegate));
        if (hip->Initialize() != Compute::ComputeResult::SUCCESS)
        {
            lastError = Widen(hip->GetLastError());
            return false;
        }
        const Compute::GpuInfo info = hip->GetGpuInfo();
        sel.backend = hip;
        sel.usedGpu = true;
        sel.kind = ComputeBackendKind::Hip;
        sel.recommendedMaxInFlight = 2;
        sel.recommendedChunkSize = RecommendedChunkSize(info.totalMemory);
        sel.statusMessage = L"Using GPU (HIP): " + Widen(info.name);
        return true;
    }

    BackendSelection SelectComputeBackend(
        ComputeBackendKind preferred,
        const std::function<std::shared_ptr<Compute::IComputeBackend>()>& makeCpu)
    {
        BackendSelection sel;
        std::wstring lastError;
        std::shared_ptr<Compute::IComputeBackend> cpu = makeCpu ? makeCpu() : nullptr;
        if (preferred == ComputeBackendKind::Auto || preferred == ComputeBackendKind::Hip)
            if (TryHip(sel, lastError, cpu)) return sel;
        sel = BackendSelection{};
        sel.backend = cpu;
        sel.usedGpu = false;
        sel.kind = ComputeBackendKind::Cpu;
        sel.recommendedMaxInFlight =
            std::max<size_t>(1, static_cast<size_t>(std::thread::hardware_concurrency()) - 1);
        if (cpu)
        {
            cpu->Initialize();
            sel.statusMessage = lastError.empty() ? L"Using " + Widen(cpu->GetGpuInfo().name)
                                                  : L"No compatible GPU (" + lastError + L") - using " +
                                                        Widen(cpu->GetGpuInfo().name);
        }
        else
        {
            sel.statusMessage = L"No compatible GPU (" + lastError + L") and no CPU backend was supplied";
        }
        return sel;
    }
}}
