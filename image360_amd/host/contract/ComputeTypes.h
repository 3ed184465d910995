// The value types of the compute-backend contract, declared with the same
// names, members and enumerator values as WindowsApp.Compute/HeaderFiles/
// ComputeTypes.h:1-89 so that HipComputeBackend compiles unchanged against
// either this header (standalone build) or the reference's own (in-tree
// build: put the reference's HeaderFiles directory first on the include
// path and this directory is never read).
#pragma once

#include <cstddef>
#include <cstdint>

namespace WindowsApp { namespace Compute
{
    // Device identity (ComputeTypes.h:18-29).  The HIP backend fills the
    // compute version with the gfx major/minor and reports matrix cores.
    struct GpuInfo
    {
        int deviceId = -1;
        char name[256] = {};
        size_t totalMemory = 0;
        size_t freeMemory = 0;
        int computeMajor = 0;
        int computeMinor = 0;
        int maxThreadsPerBlock = 0;
        int multiProcessorCount = 0;
        bool hasTensorCores = false;
    };

    // Result codes (ComputeTypes.h:32-50); the values are part of the
    // contract and equal the C ABI's I360HIP_* codes.
    enum class ComputeResult
    {
        SUCCESS = 0,
        NO_GPU = 1,
        INVALID_PARAM = 2,
        OUT_OF_MEMORY = 3,
        KERNEL_LAUNCH_FAILED = 4,
        CUDA_ERROR = 5,
        NOT_SUPPORTED = 6
    };

    struct FeaturePoint
    {
        float x = 0.0f;
        float y = 0.0f;
    };

    // Integer per-tile translation (burst alignment output).
    struct TileOffset
    {
        int dx = 0;
        int dy = 0;
    };

    // Sub-pixel per-tile translation.
    struct TileOffsetF
    {
        float dx = 0.0f;
        float dy = 0.0f;
    };

    using BriefDescriptor = uint64_t[4];  // 256 descriptor bits

    struct MatchResult
    {
        int indexA = 0;
        int indexB = 0;
        int hammingDistance = 0;
    };
}}
