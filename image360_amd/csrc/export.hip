// Preview export composite (PanoramaExporter.cpp:9-24, 34-98) and the
// float render's RGB48 output (SeamBlend.cpp:527-534).
//
// The reference pastes every rendered chunk, in chunk-list order, into one
// RGB8 canvas by nearest-neighbour downscaling: destination pixel (dx, dy)
// of a chunk's placement takes source pixel (int)(dx / scale),
// (int)(dy / scale) of the chunk's RGB48 buffer, shifted right by 8.  A
// later chunk overwrites an earlier one; a destination pixel whose source
// falls outside the buffer is left as it was.
//
// GPU form: one launch writes every output byte exactly once.  The host
// bins the output into 32x32 squares and lists, per square and in chunk
// order, the chunks whose clipped placement meets it.  Each output pixel
// scans its square's list from the last entry down and takes the first
// chunk that writes it, which is the value the in-order loop leaves.  A
// thread does four consecutive pixels (12 output bytes, three aligned
// dword stores); when they share a row and a square, one scan of the list
// decides all four and their texel loads go out together.  The division is IEEE binary32 (correctly rounded), as
// the reference's.

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace i360 {

void* ctx_scratch(i360hip_ctx* ctx, int slot, size_t bytes, Status& st);

namespace {

constexpr int kSlotExport = 20;
constexpr int kSlotExportStage = 23;  // host-buffer calls: staged chunks and output
constexpr int kBin = 32;  // output square edge of the chunk lists

struct XChunk {
    int x0, y0, w, h;  // placement in the output (destX, destY, destW, destH)
    int bw, bh;        // source buffer size
    const void* p;     // RGB48 or FP32 RGB
};

// The reference's clamp and rounding of a value in 0..65535
// (SeamBlend.cpp:531-533): max(0, min(65535, v)) then u16(v + 0.5).
__device__ __forceinline__ uint32_t u16_of(float v) {
    const float a = (v < 65535.0f) ? v : 65535.0f;
    const float b = (0.0f < a) ? a : 0.0f;
    return static_cast<uint32_t>(b + 0.5f);
}

// RGB8 of source pixel (sx, sy) of a buffer of width bw.
template <bool F32>
__device__ __forceinline__ uint32_t texel_rgb8(const void* p, int bw, int sx, int sy) {
    const size_t i = (static_cast<size_t>(sy) * bw + sx) * 3;
    uint32_t r, g, b;
    if (F32) {
        const float* s = static_cast<const float*>(p) + i;
        r = u16_of(s[0] * 65535.0f);
        g = u16_of(s[1] * 65535.0f);
        b = u16_of(s[2] * 65535.0f);
    } else {
        const uint16_t* s = static_cast<const uint16_t*>(p) + i;
        r = s[0];
        g = s[1];
        b = s[2];
    }
    return (r >> 8) | ((g >> 8) << 8) | ((b >> 8) << 16);
}

// Four pixels (x .. x+3, y) of one row and one square: one scan of the
// square's list serves all four.  Each pixel keeps the first chunk (from
// the end of the list) that writes it, as export_pixel does for one pixel;
// the scan stops once all four are decided.  The texel loads are then
// independent of each other.
template <bool F32>
__device__ __forceinline__ void export_quad(int x, int y, const XChunk* __restrict__ ch,
                                            const int* __restrict__ off, const int* __restrict__ list,
                                            int bins_x, float scale, uint32_t v[4]) {
    const int bin = (y / kBin) * bins_x + x / kBin;
    const int lo = off[bin];
    const void* src[4] = {nullptr, nullptr, nullptr, nullptr};
    int bw[4] = {0, 0, 0, 0}, sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0};
    unsigned todo = 0xFu;
    for (int k = off[bin + 1] - 1; k >= lo && todo; --k) {
        const XChunk c = ch[list[k]];
        const int dy = y - c.y0;
        if (dy < 0 || dy >= c.h) continue;
        const int ty = static_cast<int>(static_cast<float>(dy) / scale);
        if (ty < 0 || ty >= c.bh) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int dx = x + j - c.x0;
            if (!((todo >> j) & 1u) || dx < 0 || dx >= c.w) continue;
            const int tx = static_cast<int>(static_cast<float>(dx) / scale);
            if (tx < 0 || tx >= c.bw) continue;
            src[j] = c.p;
            bw[j] = c.bw;
            sx[j] = tx;
            sy[j] = ty;
            todo &= ~(1u << j);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = src[j] ? texel_rgb8<F32>(src[j], bw[j], sx[j], sy[j]) : 0u;
}

template <bool F32>
__device__ __forceinline__ uint32_t export_pixel(int x, int y, const XChunk* __restrict__ ch,
                                                 const int* __restrict__ off, const int* __restrict__ list,
                                                 int bins_x, float scale) {
    const int bin = (y / kBin) * bins_x + x / kBin;
    const int lo = off[bin];
    for (int k = off[bin + 1] - 1; k >= lo; --k) {
        const XChunk c = ch[list[k]];
        const int dx = x - c.x0, dy = y - c.y0;
        if (dx < 0 || dy < 0 || dx >= c.w || dy >= c.h) continue;
        const int sx = static_cast<int>(static_cast<float>(dx) / scale);
        const int sy = static_cast<int>(static_cast<float>(dy) / scale);
        if (sx < 0 || sx >= c.bw || sy < 0 || sy >= c.bh) continue;
        return texel_rgb8<F32>(c.p, c.bw, sx, sy);
    }
    return 0u;
}

// Pixels [4q, 4q+4) of the row-major output; `n` pixels in all.  ALIGNED:
// the output is 4-byte aligned, so a full quad is three dword stores.
template <bool F32, bool ALIGNED>
__global__ void __launch_bounds__(256) export_kernel(const XChunk* __restrict__ ch, const int* __restrict__ off,
                                                     const int* __restrict__ list, int bins_x, int out_w,
                                                     float scale, size_t n, uint8_t* __restrict__ out) {
    const size_t quads = (n + 3) / 4;
    for (size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < quads;
         q += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t p0 = 4 * q;
        int y = static_cast<int>(p0 / out_w);
        int x = static_cast<int>(p0 - static_cast<size_t>(y) * out_w);
        uint32_t v[4];
        if (p0 + 4 <= n && x + 3 < out_w && x / kBin == (x + 3) / kBin) {
            export_quad<F32>(x, y, ch, off, list, bins_x, scale, v);
        } else {
            // the quad crosses a row end or a square edge: one pixel at a time
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v[j] = (p0 + j < n) ? export_pixel<F32>(x, y, ch, off, list, bins_x, scale) : 0u;
                if (++x == out_w) {
                    x = 0;
                    ++y;
                }
            }
        }
        uint8_t* d = out + 3 * p0;
        if (ALIGNED && p0 + 4 <= n) {
            uint32_t* d32 = reinterpret_cast<uint32_t*>(d);
            d32[0] = v[0] | (v[1] << 24);
            d32[1] = (v[1] >> 8) | (v[2] << 16);
            d32[2] = (v[2] >> 16) | (v[3] << 8);
        } else {
            for (int j = 0; j < 4 && p0 + j < n; ++j) {
                d[3 * j] = static_cast<uint8_t>(v[j]);
                d[3 * j + 1] = static_cast<uint8_t>(v[j] >> 8);
                d[3 * j + 2] = static_cast<uint8_t>(v[j] >> 16);
            }
        }
    }
}

// FP32 RGB in [0, 1] -> RGB48 by the reference's output rounding; four
// values per thread where the arrays allow it.
__device__ __forceinline__ uint2 rgb48_of4(const float4 v) {
    uint2 o;
    o.x = u16_of(v.x * 65535.0f) | (u16_of(v.y * 65535.0f) << 16);
    o.y = u16_of(v.z * 65535.0f) | (u16_of(v.w * 65535.0f) << 16);
    return o;
}

__global__ void __launch_bounds__(256) rgb48_kernel(const float* __restrict__ in, uint16_t* __restrict__ out,
                                                    size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t n4 = n / 4;
    const auto* in4 = reinterpret_cast<const float4*>(in);
    auto* out2 = reinterpret_cast<uint2*>(out);
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    // four independent 16-byte loads in flight per thread
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 v0 = in4[i], v1 = in4[i + stride], v2 = in4[i + 2 * stride], v3 = in4[i + 3 * stride];
        out2[i] = rgb48_of4(v0);
        out2[i + stride] = rgb48_of4(v1);
        out2[i + 2 * stride] = rgb48_of4(v2);
        out2[i + 3 * stride] = rgb48_of4(v3);
    }
    for (; i < n4; i += stride) out2[i] = rgb48_of4(in4[i]);
    for (size_t i = 4 * n4 + static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = static_cast<uint16_t>(u16_of(in[i] * 65535.0f));
}

int grid_for(const i360hip_ctx* ctx, size_t work) {
    const size_t cap = static_cast<size_t>(std::max(1, ctx->props.multiProcessorCount)) * 8;
    const size_t want = (work + 255) / 256;
    return static_cast<int>(std::max<size_t>(1, std::min(cap, want)));
}

int Fail(i360hip_ctx* ctx, int code, const char* msg) {
    set_error(ctx, msg);
    return code;
}

}  // namespace
}  // namespace i360

using namespace i360;

extern "C" {

float i360hip_export_scale(int total_width, int total_height, int max_dimension) {
    const int longest = std::max(total_width, total_height);
    if (longest <= 0) return 1.0f;
    return static_cast<float>(max_dimension) / static_cast<float>(longest);
}

void i360hip_export_placement(int x_offset, int y_offset, int width, int height, float scale, int* out) {
    if (!out) return;
    out[0] = static_cast<int>(x_offset * scale);
    out[1] = static_cast<int>(y_offset * scale);
    out[2] = std::max(1, static_cast<int>(width * scale));
    out[3] = std::max(1, static_cast<int>(height * scale));
}

int i360hip_export_size(int total_width, int total_height, int max_dimension, int* out_width, int* out_height) {
    if (!out_width || !out_height) return I360HIP_INVALID_PARAM;
    if (max_dimension <= 0 || total_width <= 0 || total_height <= 0) return I360HIP_INVALID_PARAM;
    const float scale = i360hip_export_scale(total_width, total_height, max_dimension);
    *out_width = std::max(1, static_cast<int>(total_width * scale));
    *out_height = std::max(1, static_cast<int>(total_height * scale));
    return I360HIP_SUCCESS;
}

}  // extern "C"

namespace {
int export_validate(i360hip_ctx* ctx, const i360hip_export_chunk* chunks, int num_chunks, int total_width,
                    int total_height, int max_dimension, int format, const uint8_t* rgb8_out) {
    if (max_dimension <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "max_dimension must be positive");
    if (total_width <= 0 || total_height <= 0) return Fail(ctx, I360HIP_INVALID_PARAM, "empty panorama");
    if (num_chunks < 0 || (num_chunks > 0 && !chunks)) return Fail(ctx, I360HIP_INVALID_PARAM, "bad chunk list");
    if (!rgb8_out) return Fail(ctx, I360HIP_INVALID_PARAM, "null output");
    if (format != I360HIP_PIX_U16 && format != I360HIP_PIX_F32)
        return Fail(ctx, I360HIP_INVALID_PARAM, "chunk format must be I360HIP_PIX_U16 or I360HIP_PIX_F32");
    for (int c = 0; c < num_chunks; ++c)
        if (chunks[c].pixels && (chunks[c].buf_width < 0 || chunks[c].buf_height < 0))
            return Fail(ctx, I360HIP_INVALID_PARAM, "negative buffer size");
    return I360HIP_SUCCESS;
}

// The composite on device buffers; the caller holds ctx->mu.
int export_dev_locked(i360hip_ctx* ctx, const i360hip_export_chunk* chunks, int num_chunks, int total_width,
                      int total_height, int max_dimension, int format, uint8_t* rgb8_out, hipStream_t s) {
    const float scale = i360hip_export_scale(total_width, total_height, max_dimension);
    int out_w = 0, out_h = 0;
    (void)i360hip_export_size(total_width, total_height, max_dimension, &out_w, &out_h);
    const int bins_x = (out_w + kBin - 1) / kBin, bins_y = (out_h + kBin - 1) / kBin;
    const size_t nbins = static_cast<size_t>(bins_x) * bins_y;
    const size_t n = static_cast<size_t>(out_w) * out_h;
    const bool aligned = (reinterpret_cast<uintptr_t>(rgb8_out) & 3) == 0;

    // The tables depend only on the arguments: a repeat of the last call's
    // arguments launches on the tables already on the device.
    uint64_t key = 0x9E3779B97F4A7C15ull;
    auto mix = [&key](uint64_t v) { key = (key ^ v) * 0x100000001B3ull + (key >> 29); };
    mix(static_cast<uint64_t>(num_chunks));
    mix(static_cast<uint64_t>(total_width) << 32 | static_cast<uint32_t>(total_height));
    mix(static_cast<uint64_t>(max_dimension));
    for (int c = 0; c < num_chunks; ++c) {
        const i360hip_export_chunk& k = chunks[c];
        mix(static_cast<uint64_t>(static_cast<uint32_t>(k.x_offset)) << 32 | static_cast<uint32_t>(k.y_offset));
        mix(static_cast<uint64_t>(static_cast<uint32_t>(k.width)) << 32 | static_cast<uint32_t>(k.height));
        mix(reinterpret_cast<uintptr_t>(k.pixels));
        mix(static_cast<uint64_t>(static_cast<uint32_t>(k.buf_width)) << 32 | static_cast<uint32_t>(k.buf_height));
    }
    auto launch = [&](const uint8_t* dev, size_t oO, size_t oL) -> int {
        const auto* dch = reinterpret_cast<const XChunk*>(dev);
        const auto* doff = reinterpret_cast<const int*>(dev + oO);
        const auto* dlist = reinterpret_cast<const int*>(dev + oL);
        const int grid = grid_for(ctx, (n + 3) / 4);
        {
            // algorithmic bytes: one source texel and one RGB8 pixel per output pixel
            ProfScope ps(ctx, s, "export_composite",
                         static_cast<double>(n) * ((format == I360HIP_PIX_F32 ? 12 : 6) + 3));
            if (format == I360HIP_PIX_F32) {
                if (aligned) export_kernel<true, true><<<grid, 256, 0, s>>>(dch, doff, dlist, bins_x, out_w, scale, n, rgb8_out);
                else export_kernel<true, false><<<grid, 256, 0, s>>>(dch, doff, dlist, bins_x, out_w, scale, n, rgb8_out);
            } else {
                if (aligned) export_kernel<false, true><<<grid, 256, 0, s>>>(dch, doff, dlist, bins_x, out_w, scale, n, rgb8_out);
                else export_kernel<false, false><<<grid, 256, 0, s>>>(dch, doff, dlist, bins_x, out_w, scale, n, rgb8_out);
            }
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return Fail(ctx, I360HIP_KERNEL_LAUNCH_FAILED, hipGetErrorString(e));
        return I360HIP_SUCCESS;
    };
    if (ctx->export_dev && key == ctx->export_key && static_cast<int>(ctx->bufs.size()) > kSlotExport &&
        ctx->bufs[kSlotExport].ptr == ctx->export_dev)
        return launch(static_cast<const uint8_t*>(ctx->export_dev), ctx->export_off[0], ctx->export_off[1]);
    ctx->export_dev = nullptr;

    // chunks that can write: a buffer and a placement meeting the output
    std::vector<XChunk> xc;
    std::vector<int4> rect;  // clipped bin range per kept chunk
    xc.reserve(num_chunks);
    for (int c = 0; c < num_chunks; ++c) {
        const i360hip_export_chunk& k = chunks[c];
        if (!k.pixels || k.buf_width == 0 || k.buf_height == 0) continue;
        int pl[4];
        i360hip_export_placement(k.x_offset, k.y_offset, k.width, k.height, scale, pl);
        const long long xa = std::max<long long>(pl[0], 0), ya = std::max<long long>(pl[1], 0);
        const long long xb = std::min<long long>(static_cast<long long>(pl[0]) + pl[2], out_w);
        const long long yb = std::min<long long>(static_cast<long long>(pl[1]) + pl[3], out_h);
        if (xa >= xb || ya >= yb) continue;
        xc.push_back(XChunk{pl[0], pl[1], pl[2], pl[3], k.buf_width, k.buf_height, k.pixels});
        rect.push_back(int4{static_cast<int>(xa / kBin), static_cast<int>(ya / kBin),
                            static_cast<int>((xb - 1) / kBin), static_cast<int>((yb - 1) / kBin)});
    }
    // per-square chunk lists in chunk order (CSR)
    std::vector<int> off(nbins + 1, 0);
    for (const int4& r : rect)
        for (int by = r.y; by <= r.w; ++by)
            for (int bx = r.x; bx <= r.z; ++bx) ++off[static_cast<size_t>(by) * bins_x + bx + 1];
    for (size_t b = 0; b < nbins; ++b) off[b + 1] += off[b];
    std::vector<int> list(std::max(1, off[nbins]));
    {
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (size_t i = 0; i < rect.size(); ++i)
            for (int by = rect[i].y; by <= rect[i].w; ++by)
                for (int bx = rect[i].x; bx <= rect[i].z; ++bx)
                    list[fill[static_cast<size_t>(by) * bins_x + bx]++] = static_cast<int>(i);
    }
    auto a256 = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    const size_t oC = 0, oO = a256(sizeof(XChunk) * std::max<size_t>(1, xc.size()));
    const size_t oL = oO + a256(sizeof(int) * off.size());
    const size_t bytes = oL + a256(sizeof(int) * list.size());
    std::vector<uint8_t> host(bytes, 0);
    if (!xc.empty()) std::memcpy(host.data() + oC, xc.data(), sizeof(XChunk) * xc.size());
    std::memcpy(host.data() + oO, off.data(), sizeof(int) * off.size());
    std::memcpy(host.data() + oL, list.data(), sizeof(int) * list.size());

    Status st;
    // an earlier call's kernel, on any stream, may still read the tables
    if (hipDeviceSynchronize() != hipSuccess) return Fail(ctx, I360HIP_DEVICE_ERROR, "device synchronisation failed");
    auto* dev = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotExport, bytes, st));
    if (!st.ok()) return Fail(ctx, st.code, st.msg);
    if (hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return Fail(ctx, I360HIP_DEVICE_ERROR, "export table upload failed");

    ctx->export_key = key;
    ctx->export_dev = dev;
    ctx->export_off[0] = oO;
    ctx->export_off[1] = oL;
    return launch(dev, oO, oL);
}
}  // namespace

extern "C" {

int i360hip_export_composite_dev(i360hip_ctx* ctx, const i360hip_export_chunk* chunks, int num_chunks,
                                 int total_width, int total_height, int max_dimension, int format,
                                 uint8_t* rgb8_out, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    const int v = export_validate(ctx, chunks, num_chunks, total_width, total_height, max_dimension, format, rgb8_out);
    if (v != I360HIP_SUCCESS) return v;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    return export_dev_locked(ctx, chunks, num_chunks, total_width, total_height, max_dimension, format, rgb8_out, s);
}

int i360hip_export_composite(i360hip_ctx* ctx, const i360hip_export_chunk* chunks, int num_chunks, int total_width,
                             int total_height, int max_dimension, int format, uint8_t* rgb8_out) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    const int v = export_validate(ctx, chunks, num_chunks, total_width, total_height, max_dimension, format, rgb8_out);
    if (v != I360HIP_SUCCESS) return v;
    int out_w = 0, out_h = 0;
    (void)i360hip_export_size(total_width, total_height, max_dimension, &out_w, &out_h);
    const size_t esz = format == I360HIP_PIX_F32 ? 12 : 6;
    const size_t out_bytes = static_cast<size_t>(out_w) * out_h * 3;
    auto a256 = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    // stage the chunk buffers and the output in one block of the staging slot
    std::vector<size_t> at(num_chunks, 0);
    size_t bytes = a256(out_bytes);
    for (int c = 0; c < num_chunks; ++c) {
        at[c] = bytes;
        if (chunks[c].pixels)
            bytes += a256(static_cast<size_t>(chunks[c].buf_width) * chunks[c].buf_height * esz);
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    Status st;
    // the device copies take other addresses than any earlier call's, so
    // the tables are rebuilt; wait for earlier work before reusing the slot
    if (hipDeviceSynchronize() != hipSuccess) return Fail(ctx, I360HIP_DEVICE_ERROR, "device synchronisation failed");
    auto* base = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotExportStage, bytes, st));
    if (!st.ok()) return Fail(ctx, st.code, st.msg);
    std::vector<i360hip_export_chunk> dc(chunks, chunks + num_chunks);
    for (int c = 0; c < num_chunks; ++c) {
        if (!chunks[c].pixels) continue;
        const size_t b = static_cast<size_t>(chunks[c].buf_width) * chunks[c].buf_height * esz;
        if (b && hipMemcpyAsync(base + at[c], chunks[c].pixels, b, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            return Fail(ctx, I360HIP_DEVICE_ERROR, "chunk upload failed");
        dc[c].pixels = base + at[c];
    }
    ctx->export_dev = nullptr;  // staged addresses may repeat with other contents
    const int r = export_dev_locked(ctx, dc.data(), num_chunks, total_width, total_height, max_dimension, format, base,
                                    ctx->stream);
    if (r != I360HIP_SUCCESS) return r;
    if (hipMemcpyAsync(rgb8_out, base, out_bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return Fail(ctx, I360HIP_DEVICE_ERROR, "composite download failed");
    return I360HIP_SUCCESS;
}

int i360hip_pixel_buffer_to_blob(i360hip_ctx* ctx, const uint16_t* rgb48, int width, int height, uint8_t* blob,
                                 size_t blob_size) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!blob || (!rgb48 && width > 0 && height > 0)) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    if (width < 0 || height < 0) return Fail(ctx, I360HIP_INVALID_PARAM, "Invalid dimensions.");
    const size_t payload = static_cast<size_t>(width) * height * 3 * sizeof(uint16_t);
    if (blob_size != 8 + payload) return Fail(ctx, I360HIP_INVALID_PARAM, "blob_size must be 8 + width*height*6");
    const int32_t hdr[2] = {width, height};  // StorageEngine.cpp:132-140: int32 width, int32 height, then the values
    std::memcpy(blob, hdr, sizeof(hdr));
    if (payload == 0) return I360HIP_SUCCESS;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (hipMemcpyAsync(blob + 8, rgb48, payload, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return Fail(ctx, I360HIP_DEVICE_ERROR, "pixel download failed");
    return I360HIP_SUCCESS;
}

int i360hip_pixel_buffer_from_blob(i360hip_ctx* ctx, const uint8_t* blob, size_t blob_size, uint16_t* rgb48,
                                   size_t capacity_values, int* out_width, int* out_height, size_t* out_values) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!blob || !out_width || !out_height || !out_values) return Fail(ctx, I360HIP_INVALID_PARAM, "Null pointer argument.");
    // StorageEngine.cpp:146-167: shorter than the header or an odd payload is no pixel buffer
    if (blob_size < 8) return Fail(ctx, I360HIP_INVALID_PARAM, "blob shorter than its 8-byte header");
    const size_t payload = blob_size - 8;
    if (payload % sizeof(uint16_t) != 0) return Fail(ctx, I360HIP_INVALID_PARAM, "odd payload size");
    int32_t hdr[2];
    std::memcpy(hdr, blob, sizeof(hdr));
    const size_t values = payload / sizeof(uint16_t);
    if (values > capacity_values || (values > 0 && !rgb48))
        return Fail(ctx, I360HIP_INVALID_PARAM, "payload larger than the device buffer");
    *out_width = hdr[0];
    *out_height = hdr[1];
    *out_values = values;
    if (values == 0) return I360HIP_SUCCESS;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (hipMemcpyAsync(rgb48, blob + 8, payload, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return Fail(ctx, I360HIP_DEVICE_ERROR, "pixel upload failed");
    return I360HIP_SUCCESS;
}

int i360hip_rgb48_from_render_dev(i360hip_ctx* ctx, const float* rgb, size_t num_values, uint16_t* rgb48,
                                  void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (num_values == 0) return I360HIP_SUCCESS;
    if (!rgb || !rgb48) return Fail(ctx, I360HIP_INVALID_PARAM, "null buffer");
    if ((reinterpret_cast<uintptr_t>(rgb) & 15) || (reinterpret_cast<uintptr_t>(rgb48) & 7))
        return Fail(ctx, I360HIP_INVALID_PARAM, "buffers must be 16-byte (FP32) and 8-byte (RGB48) aligned");
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    {
        ProfScope ps(ctx, s, "rgb48_from_render", static_cast<double>(num_values) * 6);
        rgb48_kernel<<<grid_for(ctx, (num_values + 3) / 4), 256, 0, s>>>(rgb, rgb48, num_values);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return Fail(ctx, I360HIP_KERNEL_LAUNCH_FAILED, hipGetErrorString(e));
    return I360HIP_SUCCESS;
}

}  // extern "C"
