/*
 * i360hip.h -- C ABI of the MI355X (gfx950) compute backend for Image360.
 *
 * Plain C: opaque context, plain pointers and sizes, int result codes.
 * Every entry point that takes host pointers is synchronous (results are
 * valid on return), mirroring the contract of
 * WindowsApp::Compute::IComputeBackend (WindowsApp.Compute/HeaderFiles/
 * IComputeBackend.h:16-169).  Entry points ending in `_dev` take device
 * pointers and a hipStream_t (passed as void*, NULL = the context's own
 * stream) and are asynchronous on that stream.
 *
 * Thread safety: a context may be shared by several host threads; calls on
 * one context are serialised by the context (the reference's Vulkan backend
 * does the same, VulkanPipeline.cpp:71-82).  `_dev` calls on distinct
 * streams of one context are serialised at the host-side enqueue only.
 *
 * The error text of the last failing call is kept per context
 * (i360hip_last_error), like IComputeBackend::GetLastError.
 */
#ifndef I360HIP_H
#define I360HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

#define I360HIP_ABI_VERSION 1

/* Result codes: the values of WindowsApp::Compute::ComputeResult
 * (WindowsApp.Compute/HeaderFiles/ComputeTypes.h:32-50). */
enum {
    I360HIP_SUCCESS = 0,
    I360HIP_NO_GPU = 1,
    I360HIP_INVALID_PARAM = 2,
    I360HIP_OUT_OF_MEMORY = 3,
    I360HIP_KERNEL_LAUNCH_FAILED = 4,
    I360HIP_DEVICE_ERROR = 5, /* ComputeResult::CUDA_ERROR, also "not initialised" */
    I360HIP_NOT_SUPPORTED = 6
};

/* Limits shared with the reference. */
#define I360HIP_MAX_STACK_INPUTS 32   /* MedianStack_Scalar.cpp:11, CpuComputeBackend.cpp:143 */
#define I360HIP_MAX_CONTRIBUTORS 32   /* RenderExecutor.cpp:15 */
#define I360HIP_DEFAULT_BANDS 5       /* SeamBlendKernels.h:10 */
#define I360HIP_MAX_BANDS 12

typedef struct i360hip_ctx i360hip_ctx;

/* Device identity, the fields of WindowsApp::Compute::GpuInfo
 * (ComputeTypes.h:18-29) plus the gfx architecture name. */
typedef struct i360hip_device_info {
    int device_id;
    char name[256];
    size_t total_memory;
    size_t free_memory;
    int compute_major;
    int compute_minor;
    int max_threads_per_block;
    int multiprocessor_count;
    int has_matrix_cores;
    char arch[64];
} i360hip_device_info;

/* ---- lifetime (IComputeBackend::Initialize/Shutdown/GetGpuInfo/GetLastError,
 *      IComputeBackend.h:21-25) ---------------------------------------- */

/* Creates a context on `device`.  Returns I360HIP_NO_GPU when no gfx950
 * device is present or the device index is out of range. */
int i360hip_create(int device, i360hip_ctx** out_ctx);
void i360hip_destroy(i360hip_ctx* ctx);
int i360hip_get_device_info(i360hip_ctx* ctx, i360hip_device_info* out);
/* Last error text of `ctx`; with ctx == NULL, the last i360hip_create error.
 * Shared by every thread calling on the context, like the reference backends'
 * m_lastError buffer behind IComputeBackend::GetLastError. */
const char* i360hip_last_error(const i360hip_ctx* ctx);
/* Error text of the calling thread's last failing call (on any context, or
 * i360hip_create); a failure on another thread cannot replace it.  Valid
 * until this thread's next failing call.  A wrapper that reads the message
 * right after a failing call (HipComputeBackend does) uses this one; the
 * reference has no per-thread form. */
const char* i360hip_thread_error(void);
/* Library build identification (gfx target, ABI version). */
const char* i360hip_version(void);

/* ---- kernel timing ------------------------------------------------------
 * While enabled, every kernel the context launches is bracketed by HIP
 * events on its launch stream.  i360hip_profile_read synchronises the
 * recorded launches and returns, per kernel name, the launch count, the
 * summed event time and the summed algorithmic bytes (inputs read once plus
 * outputs written once at their storage precision).  No counterpart in the
 * reference, which has no profiling hooks. */
typedef struct i360hip_kernel_stat {
    char name[64];
    int launches;
    double total_ms;
    double alg_bytes;
} i360hip_kernel_stat;

/* enable != 0 starts recording (and drops earlier records); 0 stops. */
int i360hip_profile_enable(i360hip_ctx* ctx, int enable);
/* Writes up to max_stats aggregated entries, sets *num_stats. */
int i360hip_profile_read(i360hip_ctx* ctx, i360hip_kernel_stat* stats, int max_stats, int* num_stats);

/* ---- IComputeBackend operations, host pointers, synchronous ------------ */

/* IComputeBackend::WarpPerspective (IComputeBackend.h:29-32); semantics of
 * WarpPerspective_Scalar.cpp:8-45: `inv_h` maps destination -> source
 * (row-major, 9 floats); the destination pixel (x,y) samples source point
 * inv_h * (x+offset_x, y+offset_y, 1); bilinear RGB48, (0,0,0) where the
 * sample is out of bounds or |w| < 1e-10.  Bit-exact with the Scalar tier. */
int i360hip_warp_perspective_u16(i360hip_ctx* ctx, const uint16_t* src, int src_w, int src_h, uint16_t* dst,
                                 int dst_w, int dst_h, const float* inv_h, int offset_x, int offset_y);

/* IComputeBackend::MedianStack (IComputeBackend.h:35-38): sigma-clipped
 * upper median of width*height*3 independent scalars across 1..32 planes
 * (MedianStackKernels.h:19-55).  Bit-exact. */
int i360hip_median_stack_u16(i360hip_ctx* ctx, const uint16_t* const* inputs, int num_inputs, uint16_t* output,
                             int width, int height, float sigma_threshold);

/* IComputeBackend::ApplyGain (IComputeBackend.h:41-42): in place,
 * u16(clamp(v*gain)); no-op when |gain-1| < 1e-6 (GainColorOps.cpp:16-26).
 * `num_values` counts u16 values (the reference's numPixels argument). */
int i360hip_apply_gain_u16(i360hip_ctx* ctx, uint16_t* data, int num_values, float gain);

/* IComputeBackend::DemosaicBayer (IComputeBackend.h:46-49): black level,
 * white balance, same-colour 3x3 bilinear, 3x4 colour matrix
 * (BayerDemosaic_Scalar.cpp:9-50).  cam_mul: 4 floats; rgb_cam: 12 floats
 * (row-major 3x4).  Bit-exact. */
int i360hip_demosaic_bayer_u16(i360hip_ctx* ctx, const uint16_t* cfa, int width, int height, uint16_t black_level,
                               const float* cam_mul, const float* rgb_cam, uint32_t filters, uint16_t* rgb_out);

/* Kernels::SeamBlend::BlendChunkContributors (SeamBlendKernels.h:91-92,
 * SeamBlend.cpp:549-587) on RGB48 warped contributors: coverage, chamfer
 * distance, ownership, pairwise DP seams, multi-band compose.  The
 * reference fixes num_bands to 5 (pass I360HIP_DEFAULT_BANDS for parity).
 * owner_out (optional, width*height int32) receives the final owner map.
 * Bit-exact with the reference CPU implementation. */
int i360hip_blend_chunk_u16(i360hip_ctx* ctx, const uint16_t* const* warped, int num_contributors, int width,
                            int height, int num_bands, uint16_t* out, int32_t* owner_out);

/* ---- GPU-resident float chunk render ---------------------------------------
 * RenderExecutor::Execute's compute (RenderExecutor.cpp:59-140: warp every
 * contributor, apply its gain, BlendChunkContributors) for many chunks at
 * once, on frames resident in device memory, in float RGB in [0,1]:
 *   - a sample is valid where WarpPerspective_Scalar.cpp:117-123 says so;
 *     coverage is carried explicitly (no "(0,0,0) means no data");
 *   - ownership = chamfer distance of the coverage + argmax + pairwise DP
 *     seams (SeamBlend.cpp:206-445), colour costs in [0,1] units and the
 *     smoothness weight 2/65535 per pixel;
 *   - multi-band compose (SeamBlend.cpp:447-546), output clamped to [0,1],
 *     0 where no contributor covers.
 * Forms (`mode`):
 *   FP64      binary64 map, storage and arithmetic (ground truth);
 *   PLAIN     binary32 inverse of the binary32 forward matrix and binary32
 *             canvas coordinates (the reference's arithmetic), FP32
 *             storage, one chunk after another;
 *   PRECISION FP64 inverse on the GPU, reduced per 64x16 tile to an FP32
 *             local offset + Jacobian + perspective row (exact projective
 *             form in tile-local coordinates), FP16 storage of warped
 *             frames and pyramid levels, FP32 accumulation, one chunk after
 *             another;
 *   BATCHED   PRECISION arithmetic, every chunk of the call in one
 *             persistent launch per stage, replayed from a HIP graph.
 *   SURROGATE BATCHED, with the learned surrogate (i360hip_surrogate_set)
 *             predicting each chunk's low band at pyramid level
 *             K = levels-3 in place of the levels K .. levels-1 of the
 *             multi-band compose; every 16x16 level-K tile whose predicted
 *             error exceeds the threshold falls back to the plain blend
 *             (chunks without such a tile skip the reduce and compose of
 *             the levels above K).  Chunks with fewer than 4 levels are
 *             blended plainly.
 * Frames: interleaved RGB (3 values per pixel), FP16, FP32 or RGB48
 * (uint16, each value v read as v * (1/65535) in binary32: the demosaic's
 * output and the reference's stored colour blobs), device memory.  Chunks name their contributors by index into `frames` through
 * contributors[first .. first+count-1] (at most 32 are used, in order).
 * No counterpart in IComputeBackend (the reference blends on the host). */
enum {
    I360HIP_PIX_F16 = 0,
    I360HIP_PIX_F32 = 1,
    I360HIP_PIX_U16 = 2 /* RGB48 */
};
enum {
    I360HIP_RENDER_FP64 = 0,
    I360HIP_RENDER_PLAIN = 1,
    I360HIP_RENDER_PRECISION = 2,
    I360HIP_RENDER_BATCHED = 3,
    I360HIP_RENDER_SURROGATE = 4
};
/* OR-ed into `mode`: the chunks' `out` receive RGB48 (uint16_t, width*height*3)
 * instead of FP32 RGB, each value u16(clamp(v * 65535) + 0.5) of the FP32
 * value, the rounding of the reference's blend output (SeamBlend.cpp:529-534)
 * and the pixel payload of a rendered_chunk_rgb48 blob
 * (RenderExecutor.cpp:138-142).  Equal to i360hip_rgb48_from_render_dev
 * applied to the FP32 output. */
#define I360HIP_RENDER_OUT_RGB48 0x100

typedef struct i360hip_frame {
    const void* pixels; /* device, width*height*3 values */
    int width;
    int height;
    int format;         /* I360HIP_PIX_F16, I360HIP_PIX_F32 or I360HIP_PIX_U16 */
    float gain;         /* applied when |gain-1| > 1e-6 (RenderExecutor.cpp:111) */
    double h[9];        /* forward homography frame -> canvas, row-major */
} i360hip_frame;

typedef struct i360hip_chunk {
    int x, y, width, height; /* canvas rectangle */
    int first, count;        /* contributor indices: contributors[first .. first+count-1] */
    float* out;              /* device FP32 RGB, width*height*3 (RGB48 with I360HIP_RENDER_OUT_RGB48) */
    int32_t* owner;          /* optional device owner map (index into the chunk's list), or NULL */
} i360hip_chunk;

int i360hip_render_chunks(i360hip_ctx* ctx, const i360hip_frame* frames, int num_frames,
                          const i360hip_chunk* chunks, int num_chunks, const int* contributors,
                          int num_bands, int mode, void* stream);

/* Launch configuration of the BATCHED (and SURROGATE) form: chunk groups
 * rendered on concurrent streams (1-8), chamfer strip width (4, 8, 16; 0 =
 * built-in choice), persistent-grid share per group (1 = every launch asks
 * for the whole chip). */
typedef struct i360hip_tuning {
    int chunk_groups;
    int chamfer_strip;
    int grid_share;
    double ms; /* autotune: time of one render with this configuration */
} i360hip_tuning;

/* Autotuner: renders the given chunks in the BATCHED form under candidate
 * configurations (a coordinate search over strip width, chunk groups and
 * grid share, each timed with HIP events as the best of two graph
 * replays), keeps the fastest in the context for later renders and
 * returns it.  Writes the chunks' outputs; synchronous.  No counterpart in
 * the reference. */
int i360hip_render_autotune(i360hip_ctx* ctx, const i360hip_frame* frames, int num_frames,
                            const i360hip_chunk* chunks, int num_chunks, const int* contributors, int num_bands,
                            i360hip_tuning* out);
/* Sets the configuration (NULL or zero fields = built-in defaults). */
int i360hip_render_set_tuning(i360hip_ctx* ctx, const i360hip_tuning* tuning);

/* Largest |source coordinate| difference (pixels) between the PRECISION
 * form's per-tile FP32 map and the FP64 map over every output pixel of the
 * given chunks and contributors; synchronous. */
int i360hip_render_map_error(i360hip_ctx* ctx, const i360hip_frame* frames, int num_frames,
                             const i360hip_chunk* chunks, int num_chunks, const int* contributors,
                             double* max_error_px);

/* ---- learned blend surrogate ----------------------------------------------
 * A 6-layer 3x3 convolutional network, 16 -> 32 -> 32 -> 32 -> 32 -> 32 -> 4
 * channels, ReLU between layers, zero padding, BF16 weights and FP32
 * activations / accumulation.  It predicts, at pyramid level K = bands-3 of
 * a chunk, the blended low band of the multi-band compose from per-pixel
 * features (the two largest-weight contributors' level-K Gaussians, the
 * seam mask, the single-level composite and two Laplacian scales of it);
 * output channels 0-2 are the residual to the composite (times 1/64), 3 the
 * predicted log10 of its own mean squared error.  No counterpart in the
 * reference (an ONNX model is planned there, docs/COMPUTATIONAL_PHOTOGRAPHY.md:
 * 328-360).
 *
 * Parameters: weights, per layer HWIO ([ky][kx][cin][cout]), layers back to
 * back (I360HIP_SURR_WEIGHTS values); biases, per layer [cout]
 * (I360HIP_SURR_BIASES floats).  Features / targets: NHWC FP32 batches of n
 * images of h x w.  All pointers are device pointers; calls are
 * asynchronous on `stream` (NULL = the context's stream). */
#define I360HIP_SURR_LAYERS 6
#define I360HIP_SURR_IN 16
#define I360HIP_SURR_HIDDEN 32
#define I360HIP_SURR_OUT 4
#define I360HIP_SURR_WEIGHTS 42624 /* 9*(16*32 + 4*32*32 + 32*4) */
#define I360HIP_SURR_BIASES 164    /* 5*32 + 4 */

/* out: n*h*w*4 floats. */
int i360hip_surrogate_forward(i360hip_ctx* ctx, const uint16_t* weights, const float* biases, const float* feat,
                              int n, int h, int w, float* out, void* stream);
/* Training loss (scalar) and its gradients for targets `target` (n*h*w*3,
 * the blended low band):
 *   L = mean_p [ 1e4 * mean_c d_c^2 + 1e-2 * (o_3 - log10(mean_c d_c^2 + 1e-10))^2 ],
 *   d = o_{0..2}/64 - (target - composite), composite = feat channels 7-9;
 * the error term's target is held constant. */
int i360hip_surrogate_gradients(i360hip_ctx* ctx, const uint16_t* weights, const float* biases, const float* feat,
                                const float* target, int n, int h, int w, float* grad_weights, float* grad_biases,
                                float* loss, void* stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected at `step` >= 1)
 * on FP32 master weights and biases; refreshes the BF16 weights.
 * adam_state: 2*(I360HIP_SURR_WEIGHTS + I360HIP_SURR_BIASES) floats (first
 * and second moments of the weights, then of the biases), zero at step 1. */
int i360hip_surrogate_train_step(i360hip_ctx* ctx, float* master_weights, float* biases, float* adam_state,
                                 uint16_t* weights, const float* feat, const float* target, int n, int h, int w,
                                 float lr, int step, float* loss, void* stream);
/* BF16 (round to nearest even) copy of FP32 master weights. */
int i360hip_surrogate_round_bf16(i360hip_ctx* ctx, const float* master_weights, uint16_t* weights, void* stream);

/* Parameters of the SURROGATE render form (device pointers, copied; the
 * threshold is a mean squared error per 16x16 level-K tile: 1e-5 is the
 * 50 dB PSNR equivalent on [0,1] RGB). */
int i360hip_surrogate_set(i360hip_ctx* ctx, const uint16_t* weights, const float* biases, float threshold_mse);
/* Per-pixel bound of the SURROGATE form's fallback test: a 16x16 level-K
 * tile also keeps the plain blend when any of its pixels' predicted squared
 * error exceeds max_sq_err (1e-6 aims at 1e-3 max abs on [0,1] RGB; the
 * prediction is the network's, not a guarantee).  0 (the default) turns the
 * test off.  No counterpart in the reference (the surrogate is this
 * backend's). */
int i360hip_surrogate_set_max_error(i360hip_ctx* ctx, float max_sq_err);
/* Level K the surrogate works at for a chunk of width x height at
 * num_bands (0 when the chunk has fewer than 4 levels) and its size. */
int i360hip_surrogate_chunk_level(int width, int height, int num_bands, int* level, int* level_width,
                                  int* level_height);
/* Training data from the plain multi-band blend (BATCHED arithmetic): for
 * every chunk with a surrogate level, in order, its level-K features
 * (level_width*level_height*16 floats into `feat`) and the blend's level-K
 * reconstruction (level_width*level_height*3 floats into `target`),
 * images back to back.  The chunks' `out` and `owner` are not written (nor
 * read).  Device pointers; asynchronous on `stream`. */
int i360hip_surrogate_features(i360hip_ctx* ctx, const i360hip_frame* frames, int num_frames,
                               const i360hip_chunk* chunks, int num_chunks, const int* contributors, int num_bands,
                               float* feat, float* target, void* stream);
/* Tiles of the last SURROGATE render and how many fell back; synchronous. */
int i360hip_surrogate_stats(i360hip_ctx* ctx, int* tiles, int* fallback_tiles);

/* ---- device-pointer forms (asynchronous on `stream`) ------------------- */

/* `inputs` is a HOST array of num_inputs DEVICE pointers; `num_scalars`
 * counts u16 values per plane. */
int i360hip_median_stack_u16_dev(i360hip_ctx* ctx, const uint16_t* const* inputs, int num_inputs, uint16_t* output,
                                 size_t num_scalars, float sigma_threshold, void* stream);
int i360hip_demosaic_bayer_u16_dev(i360hip_ctx* ctx, const uint16_t* cfa, int width, int height,
                                   uint16_t black_level, const float* cam_mul, const float* rgb_cam,
                                   uint32_t filters, uint16_t* rgb_out, void* stream);
/* FP32 RGB output: the colour-matrix value clamped to [0,65535] before the
 * u16 truncation (truncating it gives the u16 entry point's output). */
int i360hip_demosaic_bayer_f32_dev(i360hip_ctx* ctx, const uint16_t* cfa, int width, int height,
                                   uint16_t black_level, const float* cam_mul, const float* rgb_cam,
                                   uint32_t filters, float* rgb_out, void* stream);
int i360hip_warp_perspective_u16_dev(i360hip_ctx* ctx, const uint16_t* src, int src_w, int src_h, uint16_t* dst,
                                     int dst_w, int dst_h, const float* inv_h, int offset_x, int offset_y,
                                     void* stream);
int i360hip_apply_gain_u16_dev(i360hip_ctx* ctx, uint16_t* data, size_t num_values, float gain, void* stream);
/* `warped` is a HOST array of DEVICE pointers. */
int i360hip_blend_chunk_u16_dev(i360hip_ctx* ctx, const uint16_t* const* warped, int num_contributors, int width,
                                int height, int num_bands, uint16_t* out, int32_t* owner_out, void* stream);

/* ---- preview export ---------------------------------------------------------
 * PanoramaExporter (PanoramaExporter.cpp:9-24, 34-98) without the project
 * database, the codec and the file write: the composite of rendered chunks
 * into one RGB8 canvas whose longest edge is max_dimension. */

/* ComputeExportScale (PanoramaExporter.cpp:9-14): max_dimension / longest
 * edge in binary32, 1 when the longest edge is not positive. */
float i360hip_export_scale(int total_width, int total_height, int max_dimension);
/* ComputeChunkPlacement (PanoramaExporter.cpp:16-24): out[0..3] = destX,
 * destY, destW, destH, each the truncated binary32 product with `scale`;
 * destW and destH are at least 1. */
void i360hip_export_placement(int x_offset, int y_offset, int width, int height, float scale, int* out);
/* Composite size (PanoramaExporter.cpp:49-51); INVALID_PARAM where
 * ExportPreviewJpeg returns false for the sizes (max_dimension <= 0, empty
 * panorama). */
int i360hip_export_size(int total_width, int total_height, int max_dimension, int* out_width, int* out_height);

typedef struct i360hip_export_chunk {
    int x_offset, y_offset, width, height; /* ChunkModel's canvas rectangle (Types.h) */
    const void* pixels;                    /* device: the chunk's rendered buffer, or NULL when it has no
                                              render output (skipped, PanoramaExporter.cpp:55-68) */
    int buf_width, buf_height;             /* that buffer's PixelBuffer width and height */
} i360hip_export_chunk;

/* The composite loop of ExportPreviewJpeg (PanoramaExporter.cpp:53-98):
 * chunks in list order, nearest-neighbour source pixel
 * ((int)(dx / scale), (int)(dy / scale)), 16 -> 8 bits by >> 8, later
 * chunks over earlier ones, a pixel whose source lies outside the buffer
 * left as it was, 0 where nothing lands.  `format`: I360HIP_PIX_U16 for
 * RGB48 buffers (the reference's rendered_chunk_rgb48 blobs), or
 * I360HIP_PIX_F32 for this library's FP32 render output, which is first
 * rounded to RGB48 as i360hip_rgb48_from_render_dev does.  rgb8_out:
 * device, out_width * out_height * 3 bytes (i360hip_export_size), every
 * byte written.  A call whose arguments differ from the previous call's
 * waits for the device to go idle, uploads its tables and enqueues; a repeat
 * of the previous arguments only enqueues. */
int i360hip_export_composite_dev(i360hip_ctx* ctx, const i360hip_export_chunk* chunks, int num_chunks,
                                 int total_width, int total_height, int max_dimension, int format,
                                 uint8_t* rgb8_out, void* stream);
/* The same composite on host buffers, as ExportPreviewJpeg holds them (the
 * PixelBuffers read back from storage, PanoramaExporter.cpp:68): chunks'
 * `pixels` and rgb8_out are host memory.  Stages the buffers in device
 * memory, composites and copies the result back; synchronous. */
int i360hip_export_composite(i360hip_ctx* ctx, const i360hip_export_chunk* chunks, int num_chunks, int total_width,
                             int total_height, int max_dimension, int format, uint8_t* rgb8_out);

/* FP32 RGB render output in [0,1] -> the RGB48 values the reference's blend
 * stores (SeamBlend.cpp:529-534: clamp to [0,65535], then u16(v + 0.5)),
 * applied to v * 65535.  This is the pixel payload of a
 * rendered_chunk_rgb48 blob (RenderExecutor.cpp:138-142).  rgb: 16-byte
 * aligned; rgb48: 8-byte aligned; num_values counts floats. */
int i360hip_rgb48_from_render_dev(i360hip_ctx* ctx, const float* rgb, size_t num_values, uint16_t* rgb48,
                                  void* stream);

/* StorageEngine::WritePixelBuffer's framing (StorageEngine.cpp:130-144) of a
 * device RGB48 buffer: blob = int32 width, int32 height (host byte order),
 * then the width*height*3 values; blob_size must be 8 + width*height*6.  The
 * bytes then go to StorageEngine::WriteBlob under the tag
 * rendered_chunk_rgb48.  Synchronous. */
int i360hip_pixel_buffer_to_blob(i360hip_ctx* ctx, const uint16_t* rgb48, int width, int height, uint8_t* blob,
                                 size_t blob_size);
/* StorageEngine::ReadPixelBuffer (StorageEngine.cpp:146-167) into device
 * memory: INVALID_PARAM where the reference returns nullopt (shorter than
 * the header, odd payload) or when the payload exceeds capacity_values.
 * Like the reference, the payload is not checked against width*height*3;
 * out_values receives its length in values.  Synchronous. */
int i360hip_pixel_buffer_from_blob(i360hip_ctx* ctx, const uint8_t* blob, size_t blob_size, uint16_t* rgb48,
                                   size_t capacity_values, int* out_width, int* out_height, size_t* out_values);

/* ---- optimize-stage colour operations --------------------------------------
 * IComputeBackend::ComputeLabStats and ::ApplyReinhardColorTransfer
 * (IComputeBackend.h:65-73; CpuComputeBackend.cpp:234-257 for the
 * validation order and messages; GainColorOps.cpp:28-157 for the maths).
 * pow and cbrt are correctly rounded here; the reference's C library
 * versions are not, so results match the reference to within 1e-4 (Lab
 * mean and deviation) and 2 LSB per unit of transfer gain (transferred
 * RGB48, over 99 % identical) rather than bit for bit.  The statistics are deterministic on a device. */
int i360hip_compute_lab_stats(i360hip_ctx* ctx, const uint16_t* rgb, int width, int height, double* out_mean,
                              double* out_std);
int i360hip_apply_reinhard_u16(i360hip_ctx* ctx, uint16_t* rgb_in_out, int width, int height,
                               const double* src_mean, const double* src_std, const double* ref_mean,
                               const double* ref_std);
/* Device-image forms: rgb is device memory; the statistics and transfer
 * parameters are host arrays of 3 doubles.  The statistics call returns
 * after its result has reached the host; the transfer is enqueued. */
int i360hip_compute_lab_stats_dev(i360hip_ctx* ctx, const uint16_t* rgb, int width, int height, double* out_mean,
                                  double* out_std, void* stream);
int i360hip_apply_reinhard_u16_dev(i360hip_ctx* ctx, uint16_t* rgb_in_out, int width, int height,
                                   const double* src_mean, const double* src_std, const double* ref_mean,
                                   const double* ref_std, void* stream);
/* i360hip_compute_lab_stats_dev that allocates nothing: no counterpart in
 * the reference; it is for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results, bit for bit (same kernels and
 * grid).  Its scratch (the sRGB table, the block partials, the result;
 * 352 KB) is a static device array that each GPU holds from the library's
 * load until the process ends, shared by every call on that GPU: calls
 * that overlap on one GPU race (calls on one context are serialised).
 * INVALID_PARAM when the device needs more block partials than the array
 * holds (2048; MI355X needs 1024).  Returns after the result has reached
 * the host. */
int i360hip_compute_lab_stats_dev_noalloc(i360hip_ctx* ctx, const uint16_t* rgb, int width, int height,
                                          double* out_mean, double* out_std, void* stream);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* I360HIP_H */
