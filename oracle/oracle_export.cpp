// Oracle for the preview export composite and the RGB48 rounding of the
// float render output.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
// calls this file; only the test suite and the benchmark's CPU-baseline leg
// may use it, and only as the checker.
//
// Restates PanoramaExporter.cpp:9-24 (ComputeExportScale,
// ComputeChunkPlacement) and the composite loop of ExportPreviewJpeg
// (PanoramaExporter.cpp:49-98) as a literal in-order loop: chunks in list
// order, each writing its placement, later chunks over earlier ones.  Built
// with -ffp-contract=off and without -ffast-math, so each float operation is
// one IEEE binary32 operation, as in the reference.
//
// Parity pins (tests/test_oracle_export.py): the reference's own unit tests
// (PanoramaExporterTests.cpp:16-66) for scale and placement; closed forms of
// the composite (scale 1 is a copy >> 8, scale 1/2 takes every second
// pixel, scale float(1/3) checked against exact rational division rounded
// to binary32); last-chunk-wins and out-of-buffer invariants on overlapping
// chunks; the rounding rule against hand values.

#include <algorithm>
#include <cstddef>
#include <cstdint>

extern "C" {

// PanoramaExporter.cpp:9-14
float orc_export_scale(int totalWidth, int totalHeight, int maxDimension) {
    int longestEdge = totalWidth > totalHeight ? totalWidth : totalHeight;
    if (longestEdge <= 0) return 1.0f;
    return static_cast<float>(maxDimension) / static_cast<float>(longestEdge);
}

// PanoramaExporter.cpp:16-24
void orc_export_placement(int x, int y, int w, int h, float scale, int* out) {
    out[0] = static_cast<int>(x * scale);
    out[1] = static_cast<int>(y * scale);
    int dw = static_cast<int>(w * scale);
    int dh = static_cast<int>(h * scale);
    out[2] = dw < 1 ? 1 : dw;
    out[3] = dh < 1 ? 1 : dh;
}

// PanoramaExporter.cpp:49-51; returns 0 where ExportPreviewJpeg returns false.
int orc_export_size(int totalWidth, int totalHeight, int maxDimension, int* outW, int* outH) {
    if (maxDimension <= 0 || totalWidth <= 0 || totalHeight <= 0) return 0;
    float scale = orc_export_scale(totalWidth, totalHeight, maxDimension);
    int w = static_cast<int>(totalWidth * scale);
    int h = static_cast<int>(totalHeight * scale);
    *outW = w < 1 ? 1 : w;
    *outH = h < 1 ? 1 : h;
    return 1;
}

// SeamBlend.cpp:529-534 applied to v * 65535 (the float render's [0,1]
// values in the reference's 0..65535 units).
static inline uint16_t RoundU16(float v) {
    v = v * 65535.0f;
    float m = (v < 65535.0f) ? v : 65535.0f;   // std::min(65535.0f, v)
    float c = (0.0f < m) ? m : 0.0f;           // std::max(0.0f, m)
    return static_cast<uint16_t>(c + 0.5f);
}

void orc_rgb48_from_render(const float* in, long long n, uint16_t* out) {
    for (long long i = 0; i < n; ++i) out[i] = RoundU16(in[i]);
}

// PanoramaExporter.cpp:53-98.  Chunk c: rectangle rect[4c..4c+3] = x_offset,
// y_offset, width, height; buffer bufs[c] (NULL = no render output) of
// bufW[c] x bufH[c] pixels, RGB48 (isF32 = 0) or FP32 RGB in [0,1]
// (isF32 = 1, rounded to RGB48 first).  out: outW*outH*3 bytes.
int orc_export_composite(const int* rect, const void* const* bufs, const int* bufW, const int* bufH, int numChunks,
                         int totalWidth, int totalHeight, int maxDimension, int isF32, uint8_t* out) {
    int outWidth = 0, outHeight = 0;
    if (!orc_export_size(totalWidth, totalHeight, maxDimension, &outWidth, &outHeight)) return 0;
    float scale = orc_export_scale(totalWidth, totalHeight, maxDimension);
    std::fill(out, out + static_cast<size_t>(outWidth) * outHeight * 3, uint8_t(0));
    for (int c = 0; c < numChunks; ++c) {
        if (!bufs[c]) continue;
        int pl[4];
        orc_export_placement(rect[4 * c], rect[4 * c + 1], rect[4 * c + 2], rect[4 * c + 3], scale, pl);
        for (int dy = 0; dy < pl[3]; ++dy) {
            int outY = pl[1] + dy;
            if (outY < 0 || outY >= outHeight) continue;
            for (int dx = 0; dx < pl[2]; ++dx) {
                int outX = pl[0] + dx;
                if (outX < 0 || outX >= outWidth) continue;
                int srcX = static_cast<int>(dx / scale);
                int srcY = static_cast<int>(dy / scale);
                if (srcX < 0 || srcX >= bufW[c] || srcY < 0 || srcY >= bufH[c]) continue;
                size_t s = (static_cast<size_t>(srcY) * bufW[c] + srcX) * 3;
                size_t d = (static_cast<size_t>(outY) * outWidth + outX) * 3;
                for (int k = 0; k < 3; ++k) {
                    uint16_t v = isF32 ? RoundU16(static_cast<const float*>(bufs[c])[s + k])
                                       : static_cast<const uint16_t*>(bufs[c])[s + k];
                    out[d + k] = static_cast<uint8_t>(v >> 8);
                }
            }
        }
    }
    return 1;
}

}  // extern "C"
