// Internal launch interface of the HIP kernels (host side of each .hip).
#pragma once

#include "common.hpp"

namespace i360 {

struct PlaneTable {
    const uint16_t* p[I360HIP_MAX_STACK_INPUTS];
};

struct Mat3f {
    float v[9];
};

struct DemosaicParams {
    float cam_mul[4];
    float rgb_cam[12];
    uint32_t filters;
};

hipError_t launch_median_stack(const PlaneTable& planes, int count, uint16_t* out, size_t n, float thr,
                               hipStream_t s);
hipError_t launch_demosaic(const uint16_t* cfa, int w, int h, unsigned short black, const DemosaicParams& prm,
                           void* out, bool float_out, hipStream_t s);
hipError_t launch_warp_u16(const uint16_t* src, int sw, int sh, uint16_t* dst, int dw, int dh, const Mat3f& m,
                           int ox, int oy, hipStream_t s);
hipError_t launch_gain_u16(uint16_t* d, size_t n, float g, hipStream_t s);

}  // namespace i360
