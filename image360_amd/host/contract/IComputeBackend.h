// The compute-backend interface, declared with the same virtual members in
// the same order and with the same parameter lists as WindowsApp.Compute/
// HeaderFiles/IComputeBackend.h:16-169 (a drop-in implementation must
// override exactly these).  In an in-tree build the reference's header is
// found first on the include path and this one is unused.
#pragma once

#include "ComputeTypes.h"

namespace WindowsApp { namespace Compute
{
    class IComputeBackend
    {
    public:
        virtual ~IComputeBackend() = default;

        // lifetime (IComputeBackend.h:21-25)
        virtual ComputeResult Initialize() = 0;
        virtual void Shutdown() = 0;
        virtual bool IsInitialized() const = 0;
        virtual GpuInfo GetGpuInfo() const = 0;
        virtual const char* GetLastError() const = 0;

        // render / ingest operations, RGB48 host buffers (IComputeBackend.h:29-49)
        virtual ComputeResult WarpPerspective(
            const unsigned short* srcData, int srcW, int srcH,
            unsigned short* dstData, int dstW, int dstH,
            const float* homography, int offsetX, int offsetY) = 0;
        virtual ComputeResult MedianStack(
            const unsigned short** inputs, int numInputs,
            unsigned short* output, int width, int height,
            float sigmaThreshold = 2.0f) = 0;
        virtual ComputeResult ApplyGain(
            unsigned short* data, int numPixels, float gain) = 0;
        virtual ComputeResult DemosaicBayer(
            const unsigned short* cfaData, int width, int height,
            unsigned short blackLevel, const float camMul[4], const float rgbCam[3][4],
            uint32_t filters, unsigned short* rgbOut) = 0;

        // align / optimize operations (IComputeBackend.h:53-73)
        virtual ComputeResult DetectAndDescribeFeatures(
            const unsigned char* rgbImage, int width, int height,
            FeaturePoint* outPoints, BriefDescriptor* outDescriptors, int* outCount, int maxPoints) = 0;
        virtual ComputeResult MatchFeatures(
            const BriefDescriptor* descA, int countA,
            const BriefDescriptor* descB, int countB,
            MatchResult* outMatches, int* outMatchCount, int maxMatches,
            float ratioThreshold = 0.75f) = 0;
        virtual ComputeResult ComputeLabStats(
            const unsigned short* rgb, int width, int height, double outMean[3], double outStd[3]) = 0;
        virtual ComputeResult ApplyReinhardColorTransfer(
            unsigned short* rgbInOut, int width, int height,
            const double srcMean[3], const double srcStd[3],
            const double refMean[3], const double refStd[3]) = 0;

        // burst-mode operations (IComputeBackend.h:87-168)
        virtual ComputeResult BlockMatchAlign(
            const unsigned short* refData, const unsigned short* srcData,
            int width, int height, int tileSize, int searchRadius,
            TileOffset* outOffsets, int tilesX, int tilesY) = 0;
        virtual ComputeResult RobustMergeAccumulate(
            const unsigned short* const* frames, int numFrames,
            const TileOffset* const* perFrameOffsets,
            int width, int height, int tileSize, int tilesX, int tilesY,
            float sigma, unsigned short* output) = 0;
        virtual ComputeResult TileFftMerge(
            const unsigned short* const* frames, int numFrames,
            const TileOffset* const* perFrameOffsets,
            int width, int height, int tileSize, int tilesX, int tilesY,
            float noiseVariance, unsigned short* output) = 0;
        virtual ComputeResult StructureTensorKernelRegression(
            const unsigned short* const* frames, int numFrames,
            const TileOffsetF* const* perFrameOffsets,
            int width, int height, int tileSize, int tilesX, int tilesY,
            int scaleFactor, float noiseVariance, unsigned short* output) = 0;
    };
}}
