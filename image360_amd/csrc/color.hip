// Lab colour statistics and Reinhard colour transfer of the optimize stage
// (GainColorOps.cpp:28-157, behind IComputeBackend::ComputeLabStats and
// ::ApplyReinhardColorTransfer).
//
// Arithmetic: every binary32 operation of the reference is one binary32
// operation here, in the same order and uncontracted (the reference's
// translation unit is built without FMA).  pow and cbrt are evaluated in
// binary64 and rounded once to binary32, which is the correctly rounded
// binary32 result but for vanishingly rare cases.  The reference calls the
// C library's powf and cbrtf, which are not correctly rounded (glibc's
// cbrtf differs from the correctly rounded value on about one input in
// ten), so the results agree with the reference to a stated tolerance, not
// bit for bit: Lab means and deviations to 1e-4 (one binary32 ulp of a cube
// root times the 500 of the a component), transferred RGB48 values to
// 2 LSB per unit of transfer gain (the largest reference-to-source deviation
// ratio, which scales any Lab difference), with over 99 % identical.
//
// Layout: the sRGB-to-linear map of the 65536 possible 16-bit inputs is a
// table (built on the device at each call, 256 KB, L2-resident), so a pixel
// costs three table reads and three cube roots for the forward transform.
// Statistics: a fixed persistent grid accumulates binary64 sums per thread,
// reduces them per block in a fixed tree, and one block sums the block
// partials in order, so the result is deterministic on a given device.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace i360 {

void* ctx_scratch(i360hip_ctx* ctx, int slot, size_t bytes, Status& st);

namespace {

constexpr int kSlotColor = 21;     // table + block partials + result
constexpr int kSlotColorImg = 22;  // image staged by the host-pointer calls
constexpr int kLut = 65536;
constexpr int kStatsBlock = 256;

__device__ __forceinline__ float pow_rn(float x, float y) {
    return static_cast<float>(pow(static_cast<double>(x), static_cast<double>(y)));
}
__device__ __forceinline__ float cbrt_rn(float x) { return static_cast<float>(cbrt(static_cast<double>(x))); }

// sRGB transfer curves and the CIE Lab companding (GainColorOps.cpp:30-52).
__device__ __forceinline__ float srgb_to_linear(float c) {
#pragma clang fp contract(off)
    return (c <= 0.04045f) ? (c / 12.92f) : pow_rn((c + 0.055f) / 1.055f, 2.4f);
}
__device__ __forceinline__ float linear_to_srgb(float c) {
#pragma clang fp contract(off)
    return (c <= 0.0031308f) ? (c * 12.92f) : (1.055f * pow_rn(c, 1.0f / 2.4f) - 0.055f);
}
__device__ __forceinline__ float lab_f(float t) {
#pragma clang fp contract(off)
    constexpr float d = 6.0f / 29.0f;
    return (t > d * d * d) ? cbrt_rn(t) : (t / (3.0f * d * d) + 4.0f / 29.0f);
}
__device__ __forceinline__ float lab_f_inv(float t) {
#pragma clang fp contract(off)
    constexpr float d = 6.0f / 29.0f;
    return (t > d) ? (t * t * t) : (3.0f * d * d * (t - 4.0f / 29.0f));
}

// D65 white (GainColorOps.cpp:54-56)
constexpr float kXn = 0.95047f, kYn = 1.0f, kZn = 1.08883f;

// linear sRGB -> Lab (GainColorOps.cpp:58-75)
__device__ __forceinline__ void lab_of_linear(float r, float g, float b, float lab[3]) {
#pragma clang fp contract(off)
    const float X = 0.4124564f * r + 0.3575761f * g + 0.1804375f * b;
    const float Y = 0.2126729f * r + 0.7151522f * g + 0.0721750f * b;
    const float Z = 0.0193339f * r + 0.1191920f * g + 0.9503041f * b;
    const float fx = lab_f(X / kXn), fy = lab_f(Y / kYn), fz = lab_f(Z / kZn);
    lab[0] = 116.0f * fy - 16.0f;
    lab[1] = 500.0f * (fx - fy);
    lab[2] = 200.0f * (fy - fz);
}

// sRGB value of a 0..1 float, clamped and truncated to u16 (GainColorOps.cpp:99-101)
__device__ __forceinline__ uint16_t u16_trunc(float c) {
#pragma clang fp contract(off)
    const float v = c * 65535.0f;
    const float lo = (v < 0.0f) ? 0.0f : v;              // std::max(v, 0.0f)
    const float hi = (65535.0f < lo) ? 65535.0f : lo;    // std::min(lo, 65535.0f)
    return static_cast<uint16_t>(static_cast<unsigned int>(hi));
}

// Lab -> RGB48 (GainColorOps.cpp:77-102)
__device__ __forceinline__ void rgb_of_lab(const float lab[3], uint16_t* px) {
#pragma clang fp contract(off)
    const float fy = (lab[0] + 16.0f) / 116.0f;
    const float fx = fy + lab[1] / 500.0f;
    const float fz = fy - lab[2] / 200.0f;
    const float X = kXn * lab_f_inv(fx), Y = kYn * lab_f_inv(fy), Z = kZn * lab_f_inv(fz);
    const float rl = 3.2404542f * X - 1.5371385f * Y - 0.4985314f * Z;
    const float gl = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
    const float bl = 0.0556434f * X - 0.2040259f * Y + 1.0572252f * Z;
    px[0] = u16_trunc(linear_to_srgb(rl < 0.0f ? 0.0f : rl));
    px[1] = u16_trunc(linear_to_srgb(gl < 0.0f ? 0.0f : gl));
    px[2] = u16_trunc(linear_to_srgb(bl < 0.0f ? 0.0f : bl));
}

__global__ void __launch_bounds__(256) srgb_lut_kernel(float* __restrict__ lut) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < kLut) lut[v] = srgb_to_linear(static_cast<float>(v) / 65535.0f);
}

// Per-block binary64 sums of L, a, b and of their squares
// (GainColorOps.cpp:110-121): partial[6 * block + k], k = sum L, a, b, sq L, a, b.
__global__ void __launch_bounds__(kStatsBlock) lab_stats_kernel(const uint16_t* __restrict__ rgb, size_t n,
                                                               const float* __restrict__ lut,
                                                               double* __restrict__ partial) {
#pragma clang fp contract(off)
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (size_t i = static_cast<size_t>(blockIdx.x) * kStatsBlock + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kStatsBlock) {
        const uint16_t* p = rgb + 3 * i;
        float lab[3];
        lab_of_linear(lut[p[0]], lut[p[1]], lut[p[2]], lab);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = static_cast<double>(lab[c]);
            s[c] += v;
            s[3 + c] += v * v;
        }
    }
    __shared__ double red[6][kStatsBlock];
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kStatsBlock / 2; w > 0; w /= 2) {
        if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Sums the block partials in a fixed order (thread t takes blocks t,
// t + 256, ... in turn, then a fixed tree over the threads); mean and
// deviation as GainColorOps.cpp:123-129.  out[0..2] = mean, out[3..5] = std.
__global__ void __launch_bounds__(kStatsBlock) lab_stats_final_kernel(const double* __restrict__ partial,
                                                                     int blocks, size_t n,
                                                                     double* __restrict__ out) {
#pragma clang fp contract(off)
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < blocks; b += kStatsBlock)
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] += partial[6 * b + k];
    __shared__ double red[6][kStatsBlock];
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = kStatsBlock / 2; w > 0; w /= 2) {
        if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        const double np = static_cast<double>(n);
        const double mean = red[c][0] / np;
        const double var = red[3 + c][0] / np - mean * mean;
        out[c] = mean;
        out[3 + c] = sqrt(var > 0.0 ? var : 0.0);
    }
}

struct Transfer {
    double src_mean[3], ratio[3], ref_mean[3];
};

// In-place Reinhard transfer (GainColorOps.cpp:132-157).
__global__ void __launch_bounds__(256) reinhard_kernel(uint16_t* __restrict__ rgb, size_t n,
                                                       const float* __restrict__ lut, Transfer t) {
#pragma clang fp contract(off)
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        uint16_t* p = rgb + 3 * i;
        float lab[3];
        lab_of_linear(lut[p[0]], lut[p[1]], lut[p[2]], lab);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = static_cast<double>(lab[c]);
            lab[c] = static_cast<float>((v - t.src_mean[c]) * t.ratio[c] + t.ref_mean[c]);
        }
        rgb_of_lab(lab, p);
    }
}

int Fail(i360hip_ctx* ctx, int code, const char* msg) {
    set_error(ctx, msg);
    return code;
}
int Finish(i360hip_ctx* ctx, const Status& st) {
    if (!st.ok()) set_error(ctx, st.msg);
    return st.code;
}

int stats_blocks(const i360hip_ctx* ctx) { return std::max(1, ctx->props.multiProcessorCount) * 4; }

// Table, partials and result in the colour slot.
struct ColorWork {
    float* lut;
    double* partial;
    double* result;
};

// ---- scratch of the allocation-free statistics entry point --------------
// The sRGB table, the block partials (for at most kNoAllocMaxBlocks blocks)
// and the result, in one file-scope device array.  Each GPU holds its own
// copy from the library's load until the process ends; its device address
// on each GPU is looked up once (color_scratch_setup, from i360hip_create
// with that device current).  Every call rewrites all of it before reading
// it, so nothing needs clearing; calls that overlap on one GPU (two
// contexts, or streams, at once) share it and race.
constexpr int kNoAllocMaxBlocks = 2048;
constexpr size_t kNoAllocPartialOff = sizeof(float) * kLut;
constexpr size_t kNoAllocResultOff = kNoAllocPartialOff + sizeof(double) * 6 * kNoAllocMaxBlocks;
constexpr size_t kNoAllocBytes = (kNoAllocResultOff + sizeof(double) * 6 + 255) & ~static_cast<size_t>(255);

__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

ColorWork color_work(i360hip_ctx* ctx, Status& st) {
    const size_t oP = sizeof(float) * kLut;
    const size_t oR = oP + sizeof(double) * 6 * stats_blocks(ctx);
    auto* base = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotColor, oR + sizeof(double) * 6, st));
    if (!base) return ColorWork{};
    return ColorWork{reinterpret_cast<float*>(base), reinterpret_cast<double*>(base + oP),
                     reinterpret_cast<double*>(base + oR)};
}

int grid_for(const i360hip_ctx* ctx, size_t n) {
    const size_t cap = static_cast<size_t>(std::max(1, ctx->props.multiProcessorCount)) * 8;
    return static_cast<int>(std::max<size_t>(1, std::min(cap, (n + 255) / 256)));
}

// Enqueues the statistics of a device image on the table, partials and
// result of `w` and copies the six results to the host (synchronous on s).
// `timed`: bracket the launches with the context's profiling events when
// profiling is on (the allocation-free entry point takes no events).
int run_stats(i360hip_ctx* ctx, const uint16_t* rgb, size_t n, double* mean, double* sd, hipStream_t s,
              Status& st, const ColorWork& w, bool timed = true) {
    const int blocks = stats_blocks(ctx);
    ProfScope ps(timed ? ctx : nullptr, s, "lab_stats", 6.0 * n);
    srgb_lut_kernel<<<kLut / 256, 256, 0, s>>>(w.lut);
    lab_stats_kernel<<<blocks, kStatsBlock, 0, s>>>(rgb, n, w.lut, w.partial);
    lab_stats_final_kernel<<<1, kStatsBlock, 0, s>>>(w.partial, blocks, n, w.result);
    double host[6];
    I360_TRY(st, hipGetLastError());
    I360_TRY(st, hipMemcpyAsync(host, w.result, sizeof(host), hipMemcpyDeviceToHost, s));
    I360_TRY(st, hipStreamSynchronize(s));
    for (int c = 0; c < 3; ++c) {
        mean[c] = host[c];
        sd[c] = host[3 + c];
    }
    return st.code;
}

int run_reinhard(i360hip_ctx* ctx, uint16_t* rgb, size_t n, const double* sm, const double* ss, const double* rm,
                 const double* rs, hipStream_t s, Status& st) {
    ColorWork w = color_work(ctx, st);
    if (!st.ok()) return st.code;
    Transfer t{};
    for (int c = 0; c < 3; ++c) {
        const double std_src = (ss[c] > 1e-6) ? ss[c] : 1.0;  // GainColorOps.cpp:148
        t.src_mean[c] = sm[c];
        t.ratio[c] = rs[c] / std_src;
        t.ref_mean[c] = rm[c];
    }
    ProfScope ps(ctx, s, "reinhard", 12.0 * n);
    srgb_lut_kernel<<<kLut / 256, 256, 0, s>>>(w.lut);
    reinhard_kernel<<<grid_for(ctx, n), 256, 0, s>>>(rgb, n, w.lut, t);
    I360_TRY(st, hipGetLastError());
    return st.code;
}

}  // namespace

// Looks up the scratch array's address on `device` (the current device)
// once; returns false when the runtime cannot give it.
bool color_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return false;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return true;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return false;
    g_scratch_ptr[device] = static_cast<uint8_t*>(p);
    return true;
}

}  // namespace i360

using namespace i360;

extern "C" {

int i360hip_compute_lab_stats_dev(i360hip_ctx* ctx, const uint16_t* rgb, int width, int height, double* out_mean,
                                  double* out_std, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!rgb || !out_mean || !out_std) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (width <= 0 || height <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return Fail(ctx, I360HIP_DEVICE_ERROR, "hipSetDevice failed");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    const ColorWork w = color_work(ctx, st);
    if (st.ok()) run_stats(ctx, rgb, static_cast<size_t>(width) * height, out_mean, out_std, s, st, w);
    return Finish(ctx, st);
}

int i360hip_compute_lab_stats_dev_noalloc(i360hip_ctx* ctx, const uint16_t* rgb, int width, int height,
                                          double* out_mean, double* out_std, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!rgb || !out_mean || !out_std) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (width <= 0 || height <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return Fail(ctx, I360HIP_DEVICE_ERROR, "hipSetDevice failed");
    const int blocks = stats_blocks(ctx);
    if (blocks > kNoAllocMaxBlocks) {
        st.set(I360HIP_INVALID_PARAM,
               "the statistics of this device need %d block partials; the allocation-free scratch holds %d",
               blocks, kNoAllocMaxBlocks);
        return Finish(ctx, st);
    }
    uint8_t* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch_ptr[ctx->device];
    }
    if (!base) return Fail(ctx, I360HIP_DEVICE_ERROR, "the scratch array's address on this device is unknown");
    const ColorWork w{reinterpret_cast<float*>(base), reinterpret_cast<double*>(base + kNoAllocPartialOff),
                      reinterpret_cast<double*>(base + kNoAllocResultOff)};
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    run_stats(ctx, rgb, static_cast<size_t>(width) * height, out_mean, out_std, s, st, w, false);
    return Finish(ctx, st);
}

int i360hip_compute_lab_stats(i360hip_ctx* ctx, const uint16_t* rgb, int width, int height, double* out_mean,
                              double* out_std) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!rgb || !out_mean || !out_std) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (width <= 0 || height <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return Fail(ctx, I360HIP_DEVICE_ERROR, "hipSetDevice failed");
    const size_t n = static_cast<size_t>(width) * height;
    auto* d = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotColorImg, n * 6, st));
    if (!st.ok()) return Finish(ctx, st);
    if (hipMemcpyAsync(d, rgb, n * 6, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return Fail(ctx, I360HIP_DEVICE_ERROR, "image upload failed");
    const ColorWork w = color_work(ctx, st);
    if (st.ok()) run_stats(ctx, d, n, out_mean, out_std, ctx->stream, st, w);
    return Finish(ctx, st);
}

int i360hip_apply_reinhard_u16_dev(i360hip_ctx* ctx, uint16_t* rgb, int width, int height, const double* src_mean,
                                   const double* src_std, const double* ref_mean, const double* ref_std,
                                   void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!rgb || !src_mean || !src_std || !ref_mean || !ref_std)
        return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (width <= 0 || height <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return Fail(ctx, I360HIP_DEVICE_ERROR, "hipSetDevice failed");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    run_reinhard(ctx, rgb, static_cast<size_t>(width) * height, src_mean, src_std, ref_mean, ref_std, s, st);
    return Finish(ctx, st);
}

int i360hip_apply_reinhard_u16(i360hip_ctx* ctx, uint16_t* rgb, int width, int height, const double* src_mean,
                               const double* src_std, const double* ref_mean, const double* ref_std) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!rgb || !src_mean || !src_std || !ref_mean || !ref_std)
        return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (width <= 0 || height <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (hipSetDevice(ctx->device) != hipSuccess) return Fail(ctx, I360HIP_DEVICE_ERROR, "hipSetDevice failed");
    const size_t n = static_cast<size_t>(width) * height;
    auto* d = static_cast<uint16_t*>(ctx_scratch(ctx, kSlotColorImg, n * 6, st));
    if (!st.ok()) return Finish(ctx, st);
    auto run = [&]() -> int {
        I360_TRY(st, hipMemcpyAsync(d, rgb, n * 6, hipMemcpyHostToDevice, ctx->stream));
        run_reinhard(ctx, d, n, src_mean, src_std, ref_mean, ref_std, ctx->stream, st);
        if (!st.ok()) return st.code;
        I360_TRY(st, hipMemcpyAsync(rgb, d, n * 6, hipMemcpyDeviceToHost, ctx->stream));
        I360_TRY(st, hipStreamSynchronize(ctx->stream));
        return st.code;
    };
    run();
    return Finish(ctx, st);
}

}  // extern "C"
