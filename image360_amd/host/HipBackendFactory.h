// Backend selection with the HIP backend as a candidate: the parts of
// WindowsApp/WindowsApp/ComputeBackendFactory.{h,cpp} (ComputeBackendFactory.h:
// 16-44, ComputeBackendFactory.cpp:29-136) that a drop-in needs, with a
// `Hip` kind added.  The reference factory keeps its own TryCuda/TryVulkan;
// TryHip below is the function it calls for the new kind (the same shape as
// TryCuda/TryVulkan: initialise, or report why not and let the caller fall
// back to the CPU backend).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>

#include "IComputeBackend.h"

namespace WindowsApp { namespace Hip
{
    enum class ComputeBackendKind
    {
        Auto,
        Cuda,
        Vulkan,
        Cpu,
        Hip,
    };

    // The backend part of ComputeBackendSelection (ComputeBackendFactory.h:
    // 24-42); the JPEG codec pairing stays with the reference factory.
    struct BackendSelection
    {
        std::shared_ptr<Compute::IComputeBackend> backend;
        bool usedGpu = false;
        ComputeBackendKind kind = ComputeBackendKind::Cpu;
        size_t recommendedMaxInFlight = 2;
        int recommendedChunkSize = 1024;
        std::wstring statusMessage;
    };

    // Chunk edge for a GPU with totalVramBytes of memory: 4096 at >= 16 GiB,
    // 2048 at >= 8 GiB, else 1024 (ProjectManager.cpp:136-142).
    int RecommendedChunkSize(uint64_t totalVramBytes);

    // Creates and initialises a HipComputeBackend on `device`.  On success
    // fills `sel` (two tasks in flight, chunk size from device memory,
    // status "Using GPU (HIP): <name>") and returns true; otherwise stores
    // the backend's error text in lastError and returns false.
    // hostDelegate serves the align operations (may be null).
    bool TryHip(BackendSelection& sel, std::wstring& lastError,
                std::shared_ptr<Compute::IComputeBackend> hostDelegate = nullptr, int device = 0);

    // SelectComputeBackend (ComputeBackendFactory.cpp:108-136) for the kinds
    // this library can build on its own: Auto and Hip try the HIP backend;
    // when it does not initialise (and for Cpu) the selection is the backend
    // `makeCpu` returns (in the reference build, Core's CpuComputeBackend),
    // with the reason in statusMessage.  Cuda and Vulkan are the reference
    // factory's and are treated as Cpu here.
    BackendSelection SelectComputeBackend(
        ComputeBackendKind preferred,
        const std::function<std::shared_ptr<Compute::IComputeBackend>()>& makeCpu);
}}
