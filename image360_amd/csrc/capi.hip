// C ABI of the HIP backend (declared in include/i360hip.h).
//
// Host-pointer entry points follow the reference backends' call contract
// (CudaPipeline.cpp:241-265 does H2D, launch, synchronise, D2H per call):
// validate in the order CpuComputeBackend.cpp:112-193 does, stage the
// inputs into context-owned device buffers, launch on the context stream,
// copy back and synchronise.  Device-pointer entry points only validate and
// enqueue.

#include <cmath>
#include <cstdarg>
#include <cstring>

#include "blend.hpp"
#include "ops.hpp"

namespace i360 {

void Status::set(int c, const char* fmt, ...) {
    if (code != I360HIP_SUCCESS) return;
    code = c;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
}

namespace {

char g_create_error[512] = "";
std::mutex g_create_mu;
thread_local char t_error[512] = "";

enum Slot : int { kSlotA = 0, kSlotB, kSlotC, kSlotPlanes, kSlotCount };

}  // namespace

// Context-owned device buffer `slot` with at least `bytes` bytes.
void* ctx_scratch(i360hip_ctx* ctx, int slot, size_t bytes, Status& st) {
    if (static_cast<int>(ctx->bufs.size()) <= slot) ctx->bufs.resize(slot + 1);
    DevBuf& b = ctx->bufs[slot];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (b.ptr) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        st.set(CodeOf(e), "device allocation of %zu bytes failed: %s", want, HipName(e));
        return nullptr;
    }
    b.bytes = want;
    return b.ptr;
}

void set_error(i360hip_ctx* ctx, const char* msg) {
    std::snprintf(ctx->last_error, sizeof(ctx->last_error), "%s", msg);
    std::snprintf(t_error, sizeof(t_error), "%s", msg);
}

namespace {

int Finish(i360hip_ctx* ctx, const Status& st) {
    if (!st.ok()) set_error(ctx, st.msg);
    return st.code;
}

int Fail(i360hip_ctx* ctx, int code, const char* msg) {
    set_error(ctx, msg);
    return code;
}

hipStream_t StreamOf(i360hip_ctx* ctx, void* s) { return s ? static_cast<hipStream_t>(s) : ctx->stream; }

inline size_t Align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

}  // namespace
}  // namespace i360

using namespace i360;

extern "C" {

const char* i360hip_version(void) { return "i360hip abi 1 gfx950"; }

static int create_locked(int device, i360hip_ctx** out_ctx) {
    g_create_error[0] = 0;
    if (!out_ctx) {
        std::snprintf(g_create_error, sizeof(g_create_error), "Null pointer argument.");
        return I360HIP_INVALID_PARAM;
    }
    *out_ctx = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        std::snprintf(g_create_error, sizeof(g_create_error), "No HIP device: %s",
                      e != hipSuccess ? HipName(e) : "device count is 0");
        return I360HIP_NO_GPU;
    }
    if (device < 0 || device >= n) {
        std::snprintf(g_create_error, sizeof(g_create_error), "Device %d out of range (%d devices).", device, n);
        return I360HIP_NO_GPU;
    }
    hipDeviceProp_t props{};
    e = hipGetDeviceProperties(&props, device);
    if (e != hipSuccess) {
        std::snprintf(g_create_error, sizeof(g_create_error), "hipGetDeviceProperties: %s", HipName(e));
        return I360HIP_NO_GPU;
    }
    if (std::strncmp(props.gcnArchName, "gfx950", 6) != 0) {
        std::snprintf(g_create_error, sizeof(g_create_error),
                      "Device %d is %s; this library carries gfx950 code only.", device, props.gcnArchName);
        return I360HIP_NO_GPU;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        std::snprintf(g_create_error, sizeof(g_create_error), "hipSetDevice: %s", HipName(e));
        return I360HIP_NO_GPU;
    }
    if (!color_scratch_setup(device)) {
        std::snprintf(g_create_error, sizeof(g_create_error),
                      "hipGetSymbolAddress of the colour scratch array failed on device %d", device);
        return I360HIP_DEVICE_ERROR;
    }
    auto* ctx = new i360hip_ctx();
    ctx->device = device;
    ctx->props = props;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        std::snprintf(g_create_error, sizeof(g_create_error), "hipStreamCreate: %s", HipName(e));
        delete ctx;
        return I360HIP_DEVICE_ERROR;
    }
    *out_ctx = ctx;
    return I360HIP_SUCCESS;
}

void i360hip_destroy(i360hip_ctx* ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        render_release(ctx);
        for (auto& b : ctx->bufs)
            if (b.ptr) (void)hipFree(b.ptr);
        ctx->bufs.clear();
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

int i360hip_create(int device, i360hip_ctx** out_ctx) {
    std::lock_guard<std::mutex> lk(g_create_mu);
    const int rc = create_locked(device, out_ctx);
    if (rc != I360HIP_SUCCESS) std::snprintf(t_error, sizeof(t_error), "%s", g_create_error);
    return rc;
}

const char* i360hip_last_error(const i360hip_ctx* ctx) { return ctx ? ctx->last_error : g_create_error; }

const char* i360hip_thread_error(void) { return t_error; }

int i360hip_get_device_info(i360hip_ctx* ctx, i360hip_device_info* out) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!out) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::memset(out, 0, sizeof(*out));
    out->device_id = ctx->device;
    std::snprintf(out->name, sizeof(out->name), "%s", ctx->props.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", ctx->props.gcnArchName);
    size_t fr = 0, tot = 0;
    (void)hipSetDevice(ctx->device);
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
        out->free_memory = fr;
        out->total_memory = tot;
    } else {
        out->total_memory = ctx->props.totalGlobalMem;
    }
    out->compute_major = ctx->props.major;
    out->compute_minor = ctx->props.minor;
    out->max_threads_per_block = ctx->props.maxThreadsPerBlock;
    out->multiprocessor_count = ctx->props.multiProcessorCount;
    out->has_matrix_cores = 1;
    return I360HIP_SUCCESS;
}

// ---------------------------------------------------------------- warp --
int i360hip_warp_perspective_u16_dev(i360hip_ctx* ctx, const uint16_t* src, int sw, int sh, uint16_t* dst, int dw,
                                     int dh, const float* inv_h, int ox, int oy, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!src || !dst || !inv_h) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (sw <= 0 || sh <= 0 || dw <= 0 || dh <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    I360_TRY(st, hipSetDevice(ctx->device));
    Mat3f m;
    std::memcpy(m.v, inv_h, sizeof(m.v));
    ProfScope ps(ctx, StreamOf(ctx, stream), "warp_u16", 12.0 * dw * dh);
    hipError_t e = launch_warp_u16(src, sw, sh, dst, dw, dh, m, ox, oy, StreamOf(ctx, stream));
    if (e != hipSuccess) st.set(I360HIP_KERNEL_LAUNCH_FAILED, "warp launch: %s", HipName(e));
    return Finish(ctx, st);
}

int i360hip_warp_perspective_u16(i360hip_ctx* ctx, const uint16_t* src, int sw, int sh, uint16_t* dst, int dw,
                                 int dh, const float* inv_h, int ox, int oy) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!src || !dst || !inv_h) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (sw <= 0 || sh <= 0 || dw <= 0 || dh <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        const size_t sb = static_cast<size_t>(sw) * sh * 3 * 2, db = static_cast<size_t>(dw) * dh * 3 * 2;
        auto* ds = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotA, sb, st));
        auto* dd = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotB, db, st));
        if (!st.ok()) return st.code;
        I360_TRY(st, hipMemcpyAsync(ds, src, sb, hipMemcpyHostToDevice, ctx->stream));
        Mat3f m;
        std::memcpy(m.v, inv_h, sizeof(m.v));
        I360_TRY(st, launch_warp_u16(ds, sw, sh, dd, dw, dh, m, ox, oy, ctx->stream));
        I360_TRY(st, hipMemcpyAsync(dst, dd, db, hipMemcpyDeviceToHost, ctx->stream));
        I360_TRY(st, hipStreamSynchronize(ctx->stream));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

// -------------------------------------------------------------- median --
int i360hip_median_stack_u16_dev(i360hip_ctx* ctx, const uint16_t* const* inputs, int num_inputs, uint16_t* output,
                                 size_t num_scalars, float sigma, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!inputs || !output) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (num_inputs <= 0 || num_inputs > I360HIP_MAX_STACK_INPUTS)
        return Fail(ctx, I360HIP_INVALID_PARAM, "numInputs must be 1-32.");
    PlaneTable pt{};
    for (int i = 0; i < num_inputs; ++i) {
        if (!inputs[i]) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
        pt.p[i] = inputs[i];
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        ProfScope ps(ctx, StreamOf(ctx, stream), "median_stack", 2.0 * num_scalars * (num_inputs + 1));
        I360_TRY(st, launch_median_stack(pt, num_inputs, output, num_scalars, sigma, StreamOf(ctx, stream)));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

int i360hip_median_stack_u16(i360hip_ctx* ctx, const uint16_t* const* inputs, int num_inputs, uint16_t* output,
                             int width, int height, float sigma) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!inputs || !output) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (num_inputs <= 0 || num_inputs > I360HIP_MAX_STACK_INPUTS)
        return Fail(ctx, I360HIP_INVALID_PARAM, "numInputs must be 1-32.");
    if (width <= 0 || height <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    for (int i = 0; i < num_inputs; ++i)
        if (!inputs[i]) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        const size_t n = static_cast<size_t>(width) * height * 3;
        const size_t pb = Align256(n * 2);
        auto* planes = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotA, pb * num_inputs, st));
        auto* out = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotB, n * 2, st));
        if (!st.ok()) return st.code;
        PlaneTable pt{};
        for (int i = 0; i < num_inputs; ++i) {
            pt.p[i] = reinterpret_cast<const uint16_t*>(planes + pb * i);
            I360_TRY(st, hipMemcpyAsync(planes + pb * i, inputs[i], n * 2, hipMemcpyHostToDevice, ctx->stream));
        }
        I360_TRY(st, launch_median_stack(pt, num_inputs, out, n, sigma, ctx->stream));
        I360_TRY(st, hipMemcpyAsync(output, out, n * 2, hipMemcpyDeviceToHost, ctx->stream));
        I360_TRY(st, hipStreamSynchronize(ctx->stream));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

// ---------------------------------------------------------------- gain --
int i360hip_apply_gain_u16_dev(i360hip_ctx* ctx, uint16_t* data, size_t n, float gain, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!data) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (n == 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid pixel count.");
    if (std::fabs(gain - 1.0f) < 1e-6f) return I360HIP_SUCCESS;  // GainColorOps.cpp:19
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        ProfScope ps(ctx, StreamOf(ctx, stream), "gain_u16", 4.0 * n);
        I360_TRY(st, launch_gain_u16(data, n, gain, StreamOf(ctx, stream)));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

int i360hip_apply_gain_u16(i360hip_ctx* ctx, uint16_t* data, int num_values, float gain) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!data) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (num_values <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid pixel count.");
    if (std::fabs(gain - 1.0f) < 1e-6f) return I360HIP_SUCCESS;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        const size_t b = static_cast<size_t>(num_values) * 2;
        auto* d = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotA, b, st));
        if (!st.ok()) return st.code;
        I360_TRY(st, hipMemcpyAsync(d, data, b, hipMemcpyHostToDevice, ctx->stream));
        I360_TRY(st, launch_gain_u16(d, static_cast<size_t>(num_values), gain, ctx->stream));
        I360_TRY(st, hipMemcpyAsync(data, d, b, hipMemcpyDeviceToHost, ctx->stream));
        I360_TRY(st, hipStreamSynchronize(ctx->stream));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

// ------------------------------------------------------------ demosaic --
static int DemosaicDev(i360hip_ctx* ctx, const uint16_t* cfa, int w, int h, uint16_t black, const float* cam_mul,
                       const float* rgb_cam, uint32_t filters, void* out, bool f32, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!cfa || !cam_mul || !rgb_cam || !out) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (w <= 0 || h <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    DemosaicParams prm{};
    std::memcpy(prm.cam_mul, cam_mul, sizeof(prm.cam_mul));
    std::memcpy(prm.rgb_cam, rgb_cam, sizeof(prm.rgb_cam));
    prm.filters = filters;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        ProfScope ps(ctx, StreamOf(ctx, stream), f32 ? "demosaic_f32" : "demosaic_u16",
                     (f32 ? 14.0 : 8.0) * w * h);
        I360_TRY(st, launch_demosaic(cfa, w, h, black, prm, out, f32, StreamOf(ctx, stream)));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

int i360hip_demosaic_bayer_u16_dev(i360hip_ctx* ctx, const uint16_t* cfa, int w, int h, uint16_t black,
                                   const float* cam_mul, const float* rgb_cam, uint32_t filters, uint16_t* rgb_out,
                                   void* stream) {
    return DemosaicDev(ctx, cfa, w, h, black, cam_mul, rgb_cam, filters, rgb_out, false, stream);
}

int i360hip_demosaic_bayer_f32_dev(i360hip_ctx* ctx, const uint16_t* cfa, int w, int h, uint16_t black,
                                   const float* cam_mul, const float* rgb_cam, uint32_t filters, float* rgb_out,
                                   void* stream) {
    return DemosaicDev(ctx, cfa, w, h, black, cam_mul, rgb_cam, filters, rgb_out, true, stream);
}

int i360hip_demosaic_bayer_u16(i360hip_ctx* ctx, const uint16_t* cfa, int w, int h, uint16_t black,
                               const float* cam_mul, const float* rgb_cam, uint32_t filters, uint16_t* rgb_out) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!cfa || !cam_mul || !rgb_cam || !rgb_out) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (w <= 0 || h <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    DemosaicParams prm{};
    std::memcpy(prm.cam_mul, cam_mul, sizeof(prm.cam_mul));
    std::memcpy(prm.rgb_cam, rgb_cam, sizeof(prm.rgb_cam));
    prm.filters = filters;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        const size_t ib = static_cast<size_t>(w) * h * 2, ob = ib * 3;
        auto* di = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotA, ib, st));
        auto* dout = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotB, ob, st));
        if (!st.ok()) return st.code;
        I360_TRY(st, hipMemcpyAsync(di, cfa, ib, hipMemcpyHostToDevice, ctx->stream));
        I360_TRY(st, launch_demosaic(di, w, h, black, prm, dout, false, ctx->stream));
        I360_TRY(st, hipMemcpyAsync(rgb_out, dout, ob, hipMemcpyDeviceToHost, ctx->stream));
        I360_TRY(st, hipStreamSynchronize(ctx->stream));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

// --------------------------------------------------------------- blend --
int i360hip_blend_chunk_u16_dev(i360hip_ctx* ctx, const uint16_t* const* warped, int c, int w, int h,
                                int num_bands, uint16_t* out, int32_t* owner_out, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!warped || !out) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (c < 0 || c > I360HIP_MAX_CONTRIBUTORS) return Fail(ctx, I360HIP_INVALID_PARAM, "numContributors must be 0-32.");
    if (w <= 0 || h <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    if (num_bands < 1 || num_bands > I360HIP_MAX_BANDS) return Fail(ctx, I360HIP_INVALID_PARAM, "numBands must be 1-12.");
    for (int i = 0; i < c; ++i)
        if (!warped[i]) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        blend_chunk_u16(ctx, warped, c, w, h, num_bands, out, owner_out, StreamOf(ctx, stream), st);
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

int i360hip_blend_chunk_u16(i360hip_ctx* ctx, const uint16_t* const* warped, int c, int w, int h, int num_bands,
                            uint16_t* out, int32_t* owner_out) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!warped || !out) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (c < 0 || c > I360HIP_MAX_CONTRIBUTORS) return Fail(ctx, I360HIP_INVALID_PARAM, "numContributors must be 0-32.");
    if (w <= 0 || h <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    if (num_bands < 1 || num_bands > I360HIP_MAX_BANDS) return Fail(ctx, I360HIP_INVALID_PARAM, "numBands must be 1-12.");
    for (int i = 0; i < c; ++i)
        if (!warped[i]) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    auto run = [&]() -> int {
        I360_TRY(st, hipSetDevice(ctx->device));
        const size_t n = static_cast<size_t>(w) * h;
        const size_t pb = Align256(n * 6);
        auto* planes = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotA, pb * (c > 0 ? c : 1), st));
        auto* dout = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotB, n * 6, st));
        auto* down = static_cast<int32_t*>(ctx_scratch(ctx, kSlotC, n * 4, st));
        if (!st.ok()) return st.code;
        const uint16_t* dptr[I360HIP_MAX_CONTRIBUTORS];
        for (int i = 0; i < c; ++i) {
            dptr[i] = reinterpret_cast<const uint16_t*>(planes + pb * i);
            I360_TRY(st, hipMemcpyAsync(planes + pb * i, warped[i], n * 6, hipMemcpyHostToDevice, ctx->stream));
        }
        blend_chunk_u16(ctx, dptr, c, w, h, num_bands, dout, down, ctx->stream, st);
        if (!st.ok()) return st.code;
        I360_TRY(st, hipMemcpyAsync(out, dout, n * 6, hipMemcpyDeviceToHost, ctx->stream));
        if (owner_out) I360_TRY(st, hipMemcpyAsync(owner_out, down, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        I360_TRY(st, hipStreamSynchronize(ctx->stream));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

}  // extern "C"
