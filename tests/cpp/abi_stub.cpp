// CPU stand-in for the C ABI of libi360hip.so (include/i360hip.h), for the
// sanitizer runs of the C++ host layer (tests/cpp/host_concurrency_test.cpp).
// TEST INFRASTRUCTURE ONLY: it lets HipComputeBackend run on a machine with
// no GPU under ThreadSanitizer / AddressSanitizer, so that what is checked is
// the host layer's own locking and error propagation.
//
// It keeps the library's error conventions: a context serialises its calls
// with a mutex; a failing call writes the context's shared message (read
// back, unsynchronised, by i360hip_last_error, as the reference backends'
// m_lastError is) and the calling thread's own message
// (i360hip_thread_error).  A host layer that read the shared message after
// a call would race with another thread's failing call, which the
// ThreadSanitizer run reports.
//
// Behaviour: i360hip_warp_perspective_u16 copies the source's top-left
// dst_w x dst_h window (zero outside the source) and fails with a message
// naming offset_x when offset_x >= 1000; i360hip_apply_gain_u16 multiplies
// and truncates; everything else succeeds without touching its outputs or
// reports NOT_SUPPORTED.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "i360hip.h"

struct i360hip_ctx {
    std::mutex mu;
    char last_error[512] = {};
};

namespace {
thread_local char t_error[512] = "";
char g_create_error[512] = "";

int Fail(i360hip_ctx* ctx, int code, const char* msg) {
    std::snprintf(ctx->last_error, sizeof(ctx->last_error), "%s", msg);
    std::snprintf(t_error, sizeof(t_error), "%s", msg);
    return code;
}
}  // namespace

extern "C" {

int i360hip_create(int device, i360hip_ctx** out_ctx) {
    if (!out_ctx || device != 0) {
        std::snprintf(t_error, sizeof(t_error), "stub: device %d out of range", device);
        return I360HIP_NO_GPU;
    }
    *out_ctx = new i360hip_ctx();
    return I360HIP_SUCCESS;
}

void i360hip_destroy(i360hip_ctx* ctx) { delete ctx; }

int i360hip_get_device_info(i360hip_ctx* ctx, i360hip_device_info* out) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "CPU stand-in");
    std::snprintf(out->arch, sizeof(out->arch), "none");
    out->total_memory = 64ull << 30;
    out->multiprocessor_count = 1;
    return I360HIP_SUCCESS;
}

const char* i360hip_last_error(const i360hip_ctx* ctx) { return ctx ? ctx->last_error : g_create_error; }
const char* i360hip_thread_error(void) { return t_error; }

int i360hip_warp_perspective_u16(i360hip_ctx* ctx, const uint16_t* src, int src_w, int src_h, uint16_t* dst,
                                 int dst_w, int dst_h, const float*, int offset_x, int) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (offset_x >= 1000) {
        char msg[64];
        std::snprintf(msg, sizeof(msg), "stub: warp failed for offset %d", offset_x);
        return Fail(ctx, I360HIP_KERNEL_LAUNCH_FAILED, msg);
    }
    for (int y = 0; y < dst_h; ++y)
        for (int x = 0; x < dst_w; ++x)
            for (int c = 0; c < 3; ++c)
                dst[(static_cast<size_t>(y) * dst_w + x) * 3 + c] =
                    (x < src_w && y < src_h) ? src[(static_cast<size_t>(y) * src_w + x) * 3 + c] : 0;
    return I360HIP_SUCCESS;
}

int i360hip_median_stack_u16(i360hip_ctx* ctx, const uint16_t* const*, int, uint16_t*, int, int, float) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    return Fail(ctx, I360HIP_NOT_SUPPORTED, "stub: median not supported");
}

int i360hip_apply_gain_u16(i360hip_ctx* ctx, uint16_t* data, int num_values, float gain) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (int i = 0; i < num_values; ++i) {
        const float v = static_cast<float>(data[i]) * gain;
        data[i] = static_cast<uint16_t>(v < 0.0f ? 0.0f : (v > 65535.0f ? 65535.0f : v));
    }
    return I360HIP_SUCCESS;
}

int i360hip_demosaic_bayer_u16(i360hip_ctx* ctx, const uint16_t*, int, int, uint16_t, const float*, const float*,
                               uint32_t, uint16_t*) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    return Fail(ctx, I360HIP_NOT_SUPPORTED, "stub: demosaic not supported");
}

int i360hip_blend_chunk_u16(i360hip_ctx* ctx, const uint16_t* const*, int, int, int, int, uint16_t*, int32_t*) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    return Fail(ctx, I360HIP_NOT_SUPPORTED, "stub: blend not supported");
}

int i360hip_render_chunks(i360hip_ctx* ctx, const i360hip_frame*, int, const i360hip_chunk*, int, const int*, int,
                          int, void*) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    return Fail(ctx, I360HIP_NOT_SUPPORTED, "stub: render not supported");
}

int i360hip_compute_lab_stats(i360hip_ctx* ctx, const uint16_t*, int, int, double*, double*) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    return Fail(ctx, I360HIP_NOT_SUPPORTED, "stub: lab statistics not supported");
}

int i360hip_apply_reinhard_u16(i360hip_ctx* ctx, uint16_t*, int, int, const double*, const double*, const double*,
                               const double*) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    return Fail(ctx, I360HIP_NOT_SUPPORTED, "stub: colour transfer not supported");
}

}  // extern "C"
