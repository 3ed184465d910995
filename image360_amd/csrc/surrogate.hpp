// Learned blend surrogate: a 6-layer 3x3 convolutional network
// (16 -> 32 -> 32 -> 32 -> 32 -> 32 -> 4 channels, ReLU between layers,
// zero padding) with BF16 weights and FP32 activations / accumulation.
// No counterpart in the reference, which plans an ONNX model for this role
// (docs/COMPUTATIONAL_PHOTOGRAPHY.md:328-360) but has none.
//
// Layouts
//   activations  NHWC FP32: act[((n*h + y)*w + x)*C + c]
//   weights      per layer HWIO, layers back to back:
//                W_j[((ky*3 + kx)*cin + ci)*cout + co]
//   biases       per layer [cout], layers back to back
//
// Output channels: 0-2 the predicted low-band residual (scaled by
// kSurrResScale), 3 the predicted log10 of its own mean squared error.
#pragma once

#include "common.hpp"

namespace i360 {

constexpr int kSurrLayers = I360HIP_SURR_LAYERS;
constexpr int kSurrIn = I360HIP_SURR_IN;
constexpr int kSurrHidden = I360HIP_SURR_HIDDEN;
constexpr int kSurrOut = I360HIP_SURR_OUT;
constexpr float kSurrResScale = 1.0f / 64.0f;  // residual = scale * raw output
constexpr float kSurrLossScale = 1.0e4f;       // weight of the residual MSE
constexpr float kSurrErrWeight = 1.0e-2f;      // weight of the error-head loss
constexpr float kSurrErrFloor = 1.0e-10f;      // log10(err + floor)

__host__ __device__ constexpr int surr_cin(int j) { return j == 0 ? kSurrIn : kSurrHidden; }
__host__ __device__ constexpr int surr_cout(int j) { return j == kSurrLayers - 1 ? kSurrOut : kSurrHidden; }
__host__ __device__ constexpr int surr_woff(int j) {
    int o = 0;
    for (int k = 0; k < j; ++k) o += 9 * surr_cin(k) * surr_cout(k);
    return o;
}
__host__ __device__ constexpr int surr_boff(int j) {
    int o = 0;
    for (int k = 0; k < j; ++k) o += surr_cout(k);
    return o;
}
static_assert(surr_woff(kSurrLayers) == I360HIP_SURR_WEIGHTS, "weight count");
static_assert(surr_boff(kSurrLayers) == I360HIP_SURR_BIASES, "bias count");

// One image of a surrogate launch (pix0 = its first pixel in every NHWC
// buffer; tag = caller's id, the chunk index in the render) and one 16x16
// tile of an image.
struct SurrImg {
    int pix0, w, h, tag;
};
struct SurrTile {
    int img, tx, ty, pad;
};

// Runs the network over the listed tiles (device tables); tmp0 / tmp1:
// ping-pong buffers of 32 channels per pixel.
hipError_t surr_forward_tables(const SurrImg* imgs, const SurrTile* tiles, int ntiles, const uint16_t* w_bf16,
                               const float* bias, const float* feat, float* out, float* tmp0, float* tmp1,
                               hipStream_t s);

}  // namespace i360
