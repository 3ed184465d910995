// Learned blend surrogate: forward, backward and Adam kernels of the
// 6-layer 3x3 convolutional network declared in surrogate.hpp, and the C
// ABI entry points that train and run it.
//
// Convolution: an implicit GEMM on the matrix cores, one block per 16x16
// output tile.  M = the tile's pixels, N = output channels, K = (tap,
// input channel); FP32 MFMA (v_mfma_f32_16x16x4f32), so activations and
// products stay FP32.  The input tile (18x18 with halo, zero outside the
// image) and the BF16 weights (widened to FP32 by a 16-bit shift) are
// staged in LDS; each wavefront accumulates four output rows (four 16-pixel
// M blocks) over all output channels.  The grid is persistent: a block
// stages the weights once and loads the next tile's inputs into registers
// while the current tile's MFMAs run.  The backward data pass is the same
// kernel on the ReLU-masked output gradient with the weights flipped and
// transposed; the weight gradient is a per-tile partial sum (LDS-staged
// input and output-gradient tiles) reduced across tiles in a fixed order,
// so training is deterministic.

#include <atomic>
#include <cmath>
#include <vector>

#include "surrogate.hpp"

namespace i360 {

namespace {

constexpr int kSlotSurrWork = 16;
constexpr int kSlotSurrTables = 17;

constexpr int kT = 16;        // output tile edge
constexpr int kTH = kT + 2;   // tile with halo

using SImg = SurrImg;
using STile = SurrTile;

__device__ __forceinline__ float bf16_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t u) { return __uint_as_float(u & 0xFFFF0000u); }

__device__ __forceinline__ uint16_t to_bf16(float f) {
    // round to nearest even (finite values)
    const uint32_t u = __float_as_uint(f);
    return static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Stages rows [y0-1, y0+rows+1) x cols [x0-1, x0+17) of an NHWC image into
// LDS as [row][col][C]; outside the image (and outside the mask, when given)
// the value is 0.
template <int C>
__device__ __forceinline__ void stage_tile(float* lds, const float* __restrict__ in, const float* __restrict__ mask,
                                           const SImg& im, int x0, int y0, int rows) {
    constexpr int C4 = C / 4;
    const int n = (rows + 2) * kTH * C4;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int c4 = i % C4, pc = i / C4;
        const int col = pc % kTH, row = pc / kTH;
        const int x = x0 - 1 + col, y = y0 - 1 + row;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (x >= 0 && y >= 0 && x < im.w && y < im.h) {
            const size_t p = static_cast<size_t>(im.pix0) + static_cast<size_t>(y) * im.w + x;
            v = *reinterpret_cast<const float4*>(in + p * C + 4 * c4);
            if (mask) {
                const float4 m = *reinterpret_cast<const float4*>(mask + p * C + 4 * c4);
                v.x = m.x > 0.f ? v.x : 0.f;
                v.y = m.y > 0.f ? v.y : 0.f;
                v.z = m.z > 0.f ? v.z : 0.f;
                v.w = m.w > 0.f ? v.w : 0.f;
            }
        }
        *reinterpret_cast<float4*>(lds + static_cast<size_t>(pc) * C + 4 * c4) = v;
    }
}

// out = act(conv3x3(in) + bias) over the listed tiles.
//   RELU: max(0, .) on the output;  BIAS: add the bias.
//   mask: when non-null the input is multiplied by (mask > 0) (backward
//   data through a ReLU layer whose output is `mask`).
// MFMA 16x16x4 (FP32) lane layout: A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, D[i][j] in lane j + 16 (i / 4), element i % 4.  Here i is a
// pixel of one output row, j an output channel of a 16-channel block and k
// one of four consecutive input channels of one tap.
typedef float surr_f4 __attribute__((ext_vector_type(4)));
template <int CIN, int COUT, bool BIAS, bool RELU>
__global__ void __launch_bounds__(256) surr_conv_kernel(const SImg* __restrict__ imgs, const STile* __restrict__ tiles,
                                                        int ntiles, const float* __restrict__ in,
                                                        const float* __restrict__ mask, float* __restrict__ out,
                                                        const uint32_t* __restrict__ wpk,
                                                        const float* __restrict__ bias) {
    static_assert(CIN % 4 == 0, "four input channels per MFMA step");
    constexpr int CP = CIN + 2;             // LDS channel pitch: the 16 pixels of a read spread over the banks
    constexpr int NB = (COUT + 15) / 16;    // 16-channel output blocks
    constexpr int NP = NB * 16;             // output channels, padded
    constexpr int NV = kTH * kTH * (CIN / 4);  // float4 groups of one input tile
    constexpr int PER = (NV + 255) / 256;      // of them per thread
    __shared__ float xs[kTH * kTH * CP];
    __shared__ uint16_t ws[9 * CIN * NP];   // BF16 weights [tap * CIN + ci][co], zero for co >= COUT
    // weights, once per block: BF16 pairs wpk[(tap * CIN + ci) * COUT / 2 + co / 2], low half = even co
    const auto* w16 = reinterpret_cast<const uint16_t*>(wpk);
    for (int i = threadIdx.x; i < 9 * CIN * NP; i += blockDim.x) {
        const int kk = i / NP, n = i % NP;
        ws[i] = n < COUT ? w16[kk * COUT + n] : uint16_t(0);
    }
    // input rows [y0-1, y0+17) x cols [x0-1, x0+17) of tile t into registers;
    // 0 outside the image and the mask
    float4 pre[PER];
    auto fetch = [&](int t) {
        const STile tl = tiles[t];
        const SImg im = imgs[tl.img];
        const int x0 = tl.tx * kT, y0 = tl.ty * kT;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = threadIdx.x + 256 * j;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            const int c4 = i % (CIN / 4), pc = i / (CIN / 4);
            const int x = x0 - 1 + pc % kTH, y = y0 - 1 + pc / kTH;
            if (i < NV && x >= 0 && y >= 0 && x < im.w && y < im.h) {
                const size_t p = static_cast<size_t>(im.pix0) + static_cast<size_t>(y) * im.w + x;
                v = *reinterpret_cast<const float4*>(in + p * CIN + 4 * c4);
                if (mask) {
                    const float4 m = *reinterpret_cast<const float4*>(mask + p * CIN + 4 * c4);
                    v.x = m.x > 0.f ? v.x : 0.f;
                    v.y = m.y > 0.f ? v.y : 0.f;
                    v.z = m.z > 0.f ? v.z : 0.f;
                    v.w = m.w > 0.f ? v.w : 0.f;
                }
            }
            pre[j] = v;
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = threadIdx.x + 256 * j;
            if (i >= NV) continue;
            float* d = xs + (i / (CIN / 4)) * CP + 4 * (i % (CIN / 4));  // 8-byte aligned (CP even)
            *reinterpret_cast<float2*>(d) = make_float2(pre[j].x, pre[j].y);
            *reinterpret_cast<float2*>(d + 2) = make_float2(pre[j].z, pre[j].w);
        }
    };
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    int t = blockIdx.x;
    if (t < ntiles) {
        fetch(t);
        commit();
    }
    __syncthreads();
    // persistent: the next tile's inputs are in flight while this tile's MFMAs run
    for (; t < ntiles; t += gridDim.x) {
        const int tn = t + static_cast<int>(gridDim.x);
        if (tn < ntiles) fetch(tn);
        surr_f4 acc[4][NB];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[m][nb] = surr_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap % 3;
            // this lane's A column: pixel li of row (4 wave + m), tap (ky, kx)
            const float* xa = xs + ((4 * wave + ky) * kTH + li + kx) * CP + lk;
            const uint16_t* wb = ws + (tap * CIN + lk) * NP + li;
#pragma unroll
            for (int c0 = 0; c0 < CIN; c0 += 4) {
                float bv[NB];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) bv[nb] = __uint_as_float(static_cast<uint32_t>(wb[c0 * NP + nb * 16]) << 16);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float av = xa[m * kTH * CP + c0];
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[m][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[nb], acc[m][nb], 0, 0, 0);
                }
            }
        }
        const STile tl = tiles[t];
        const SImg im = imgs[tl.img];
        const int x0 = tl.tx * kT, y0 = tl.ty * kT;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int y = y0 + 4 * wave + m;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int co = nb * 16 + li;
                if (co >= COUT || y >= im.h) continue;
                const float bc = BIAS ? bias[co] : 0.0f;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int x = x0 + 4 * lk + r;
                    if (x >= im.w) continue;
                    float v = acc[m][nb][r] + bc;
                    if (RELU) v = fmaxf(v, 0.f);
                    out[(static_cast<size_t>(im.pix0) + static_cast<size_t>(y) * im.w + x) * COUT + co] = v;
                }
            }
        }
        __syncthreads();  // every wave is done reading this tile from LDS
        if (tn < ntiles) commit();
        __syncthreads();
    }
}

// Per-tile partial weight and bias gradients of one layer:
//   dW[k][ci][co] = sum_p dZ[p][co] * X[p + off_k][ci],  db[co] = sum_p dZ[p][co]
// with dZ = dY masked by (Y > 0) when Y is given.  The 16x16 tile is
// processed as two 16x8 halves to keep LDS under 64 KB.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) surr_wgrad_kernel(const SImg* __restrict__ imgs, const STile* __restrict__ tiles,
                                                         const float* __restrict__ X, const float* __restrict__ dY,
                                                         const float* __restrict__ Y, float* __restrict__ partial) {
    constexpr int HR = kT / 2;
    constexpr int PAIRS = CIN * COUT;
    constexpr int G = (PAIRS + 255) / 256;
    __shared__ float xs[(HR + 2) * kTH * CIN];
    __shared__ float dz[HR * kT * COUT];
    const STile tl = tiles[blockIdx.x];
    const SImg im = imgs[tl.img];
    const int x0 = tl.tx * kT, y0 = tl.ty * kT;
    float acc[G][9];
    float bacc = 0.f;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[g][k] = 0.f;
    for (int half = 0; half < 2; ++half) {
        const int yh = y0 + half * HR;
        stage_tile<CIN>(xs, X, nullptr, im, x0, yh, HR);
        for (int i = threadIdx.x; i < HR * kT * COUT; i += blockDim.x) {
            const int c = i % COUT, pp = i / COUT;
            const int x = x0 + pp % kT, y = yh + pp / kT;
            float v = 0.f;
            if (x < im.w && y < im.h) {
                const size_t p = static_cast<size_t>(im.pix0) + static_cast<size_t>(y) * im.w + x;
                v = dY[p * COUT + c];
                if (Y && !(Y[p * COUT + c] > 0.f)) v = 0.f;
            }
            dz[i] = v;
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int q = threadIdx.x + 256 * g;
            if (q >= PAIRS) continue;
            const int co = q % COUT, ci = q / COUT;
            for (int pp = 0; pp < HR * kT; ++pp) {
                const float d = dz[pp * COUT + co];
                const int px = pp % kT, py = pp / kT;
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    acc[g][k] = fmaf(d, xs[((py + k / 3) * kTH + (px + k % 3)) * CIN + ci], acc[g][k]);
            }
        }
        if (threadIdx.x < COUT)
            for (int pp = 0; pp < HR * kT; ++pp) bacc += dz[pp * COUT + threadIdx.x];
        __syncthreads();
    }
    float* part = partial + static_cast<size_t>(blockIdx.x) * (9 * PAIRS + COUT);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int q = threadIdx.x + 256 * g;
        if (q >= PAIRS) continue;
        const int co = q % COUT, ci = q / COUT;
#pragma unroll
        for (int k = 0; k < 9; ++k) part[(k * CIN + ci) * COUT + co] = acc[g][k];
    }
    if (threadIdx.x < COUT) part[9 * PAIRS + threadIdx.x] = bacc;
}

// out[i] = sum over t of partial[t*stride + i], t ascending.
__global__ void surr_reduce_kernel(const float* __restrict__ partial, int nparts, int stride, int count,
                                   float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int t = 0; t < nparts; ++t) s += partial[static_cast<size_t>(t) * stride + i];
    out[i] = s;
}

// Flipped, transposed BF16 weights for the backward data pass:
// wT[((8-k)*COUT + co)*CIN + ci] = w[(k*CIN + ci)*COUT + co].
__global__ void surr_transpose_kernel(const uint16_t* __restrict__ w, uint16_t* __restrict__ wT, int cin, int cout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * cin * cout) return;
    const int co = i % cout, ci = (i / cout) % cin, k = i / (cin * cout);
    wT[((8 - k) * cout + co) * cin + ci] = w[i];
}

// Loss of the last layer's output and its gradient (see surrogate.hpp):
//   residual d = kSurrResScale*o[0..2] - (target - base), base = feat[7..9]
//   L = (1/P) sum_p [ kSurrLossScale * mean_c d_c^2 + kSurrErrWeight * (o[3] - log10(mean_c d_c^2 + floor))^2 ]
// One partial loss per block.
__global__ void __launch_bounds__(256) surr_loss_kernel(const float* __restrict__ o, const float* __restrict__ feat,
                                                        const float* __restrict__ target, int npix,
                                                        float* __restrict__ dO, float* __restrict__ partial) {
    __shared__ float red[256];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (p < npix) {
        const float inv = 1.0f / static_cast<float>(npix);
        float d[3], e2 = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            d[c] = kSurrResScale * o[4 * p + c] - (target[3 * p + c] - feat[kSurrIn * p + 7 + c]);
            e2 += d[c] * d[c];
        }
        e2 *= (1.0f / 3.0f);
        const float le = log10f(e2 + kSurrErrFloor);
        const float de = o[4 * p + 3] - le;
        l = kSurrLossScale * e2 + kSurrErrWeight * de * de;
#pragma unroll
        for (int c = 0; c < 3; ++c) dO[4 * p + c] = inv * kSurrLossScale * (2.0f / 3.0f) * d[c] * kSurrResScale;
        dO[4 * p + 3] = inv * kSurrErrWeight * 2.0f * de;
        l *= inv;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam on the FP32 master copy; refreshes the BF16 copy (weights only).
__global__ void surr_adam_kernel(float* __restrict__ param, float* __restrict__ m, float* __restrict__ v,
                                 const float* __restrict__ grad, uint16_t* __restrict__ bf, int n, float lr,
                                 float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = param[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    param[i] = p;
    if (bf) bf[i] = to_bf16(p);
}

__global__ void surr_to_bf16_kernel(const float* __restrict__ w, uint16_t* __restrict__ bf, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bf[i] = to_bf16(w[i]);
}

inline size_t A256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Tables for n equal images of h x w.
struct Batch {
    std::vector<SImg> imgs;
    std::vector<STile> tiles;
};

Batch make_batch(int n, int h, int w) {
    Batch b;
    const int tx = (w + kT - 1) / kT, ty = (h + kT - 1) / kT;
    for (int i = 0; i < n; ++i) {
        b.imgs.push_back(SImg{i * w * h, w, h, 0});
        for (int y = 0; y < ty; ++y)
            for (int x = 0; x < tx; ++x) b.tiles.push_back(STile{i, x, y, 0});
    }
    return b;
}

template <int CIN, int COUT, bool BIAS, bool RELU>
void conv(const SImg* imgs, const STile* tiles, int ntiles, const float* in, const float* mask, float* out,
          const uint16_t* w, const float* bias, hipStream_t s) {
    if (ntiles <= 0) return;
    // persistent grid: as many blocks as can be resident at once (LDS and
    // registers), per CU; a block that had to wait would run as a tail
    static std::atomic<int> resident{0};  // per instantiation; the same on every call of a process
    if (resident.load(std::memory_order_relaxed) == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, surr_conv_kernel<CIN, COUT, BIAS, RELU>, 256, 0) !=
                hipSuccess || per_cu <= 0)
            per_cu = 1;
        resident.store(cus * per_cu, std::memory_order_relaxed);
    }
    const int grid = std::min(ntiles, resident.load(std::memory_order_relaxed));
    surr_conv_kernel<CIN, COUT, BIAS, RELU><<<grid, 256, 0, s>>>(imgs, tiles, ntiles, in, mask, out,
                                                                 reinterpret_cast<const uint32_t*>(w), bias);
}

// Forward through all layers; act[j] receives layer j's output.
void forward_layers(const SImg* imgs, const STile* tiles, int ntiles, const float* feat, float* const* act,
                    const uint16_t* w, const float* b, hipStream_t s) {
    conv<kSurrIn, kSurrHidden, true, true>(imgs, tiles, ntiles, feat, nullptr, act[0], w + surr_woff(0), b + surr_boff(0), s);
    for (int j = 1; j < kSurrLayers - 1; ++j)
        conv<kSurrHidden, kSurrHidden, true, true>(imgs, tiles, ntiles, act[j - 1], nullptr, act[j], w + surr_woff(j),
                                                   b + surr_boff(j), s);
    conv<kSurrHidden, kSurrOut, true, false>(imgs, tiles, ntiles, act[kSurrLayers - 2], nullptr,
                                             act[kSurrLayers - 1], w + surr_woff(kSurrLayers - 1),
                                             b + surr_boff(kSurrLayers - 1), s);
}

// Workspace of a training / gradient call on n images of h x w.
struct TrainWork {
    SImg* imgs;
    STile* tiles;
    float* act[kSurrLayers];
    float* g0;  // gradient ping-pong, 32 channels
    float* g1;
    uint16_t* wT;
    float* partial;
    float* lpart;
    float* grad;  // weights then biases
    float* loss;
    int ntiles, npix, nlblocks;
};

bool train_work(i360hip_ctx* ctx, int n, int h, int w, TrainWork& tw, Status& st, hipStream_t s) {
    const Batch b = make_batch(n, h, w);
    tw.ntiles = static_cast<int>(b.tiles.size());
    tw.npix = n * h * w;
    tw.nlblocks = (tw.npix + 255) / 256;
    const size_t P = static_cast<size_t>(tw.npix);
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += A256(bytes); return o; };
    size_t oAct[kSurrLayers];
    for (int j = 0; j < kSurrLayers; ++j) oAct[j] = take(4 * P * surr_cout(j));
    const size_t oG0 = take(4 * P * kSurrHidden), oG1 = take(4 * P * kSurrHidden);
    const size_t oWT = take(2 * 9 * kSurrHidden * kSurrHidden);
    const size_t oPart = take(4 * static_cast<size_t>(tw.ntiles) * (9 * kSurrHidden * kSurrHidden + kSurrHidden));
    const size_t oLp = take(4 * static_cast<size_t>(tw.nlblocks));
    const size_t oGrad = take(4 * static_cast<size_t>(I360HIP_SURR_WEIGHTS + I360HIP_SURR_BIASES));
    const size_t oLoss = take(4);
    auto* base = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotSurrWork, off, st));
    if (!base) return false;
    const size_t tb = A256(sizeof(SImg) * b.imgs.size()) + sizeof(STile) * b.tiles.size();
    auto* tab = static_cast<uint8_t*>(ctx_scratch(ctx, kSlotSurrTables, tb, st));
    if (!tab) return false;
    tw.imgs = reinterpret_cast<SImg*>(tab);
    tw.tiles = reinterpret_cast<STile*>(tab + A256(sizeof(SImg) * b.imgs.size()));
    // earlier calls on this stream may still read the tables: wait for them,
    // upload, and wait for the upload (the host copy is a local vector)
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpyAsync(tw.imgs, b.imgs.data(), sizeof(SImg) * b.imgs.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(tw.tiles, b.tiles.data(), sizeof(STile) * b.tiles.size(), hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        st.set(I360HIP_DEVICE_ERROR, "surrogate table upload failed");
        return false;
    }
    for (int j = 0; j < kSurrLayers; ++j) tw.act[j] = reinterpret_cast<float*>(base + oAct[j]);
    tw.g0 = reinterpret_cast<float*>(base + oG0);
    tw.g1 = reinterpret_cast<float*>(base + oG1);
    tw.wT = reinterpret_cast<uint16_t*>(base + oWT);
    tw.partial = reinterpret_cast<float*>(base + oPart);
    tw.lpart = reinterpret_cast<float*>(base + oLp);
    tw.grad = reinterpret_cast<float*>(base + oGrad);
    tw.loss = reinterpret_cast<float*>(base + oLoss);
    return true;
}

template <int CIN, int COUT>
void wgrad(const TrainWork& tw, const float* X, const float* dY, const float* Y, float* gw, float* gb, hipStream_t s) {
    surr_wgrad_kernel<CIN, COUT><<<tw.ntiles, 256, 0, s>>>(tw.imgs, tw.tiles, X, dY, Y, tw.partial);
    constexpr int nw = 9 * CIN * COUT, stride = nw + COUT;
    surr_reduce_kernel<<<(nw + 255) / 256, 256, 0, s>>>(tw.partial, tw.ntiles, stride, nw, gw);
    surr_reduce_kernel<<<1, 256, 0, s>>>(tw.partial + nw, tw.ntiles, stride, COUT, gb);
}

// dX (CIN channels) of a layer from dY (COUT channels), masked by the
// layer's ReLU output Y when given.
template <int CIN, int COUT>
void dgrad(const TrainWork& tw, const float* dY, const float* Y, float* dX, const uint16_t* w, hipStream_t s) {
    constexpr int n = 9 * CIN * COUT;
    surr_transpose_kernel<<<(n + 255) / 256, 256, 0, s>>>(w, tw.wT, CIN, COUT);
    conv<COUT, CIN, false, false>(tw.imgs, tw.tiles, tw.ntiles, dY, Y, dX, tw.wT, nullptr, s);
}

// Loss into tw.loss and all gradients into tw.grad (weights, then biases).
void gradients(const TrainWork& tw, const uint16_t* w, const float* b, const float* feat, const float* target,
               hipStream_t s) {
    constexpr int L = kSurrLayers;
    forward_layers(tw.imgs, tw.tiles, tw.ntiles, feat, tw.act, w, b, s);
    surr_loss_kernel<<<tw.nlblocks, 256, 0, s>>>(tw.act[L - 1], feat, target, tw.npix, tw.g0, tw.lpart);
    surr_reduce_kernel<<<1, 256, 0, s>>>(tw.lpart, tw.nlblocks, 1, 1, tw.loss);
    float* gw = tw.grad;
    float* gb = tw.grad + I360HIP_SURR_WEIGHTS;
    float* gcur = tw.g0;  // gradient wrt the current layer's output
    float* gnext = tw.g1;
    wgrad<kSurrHidden, kSurrOut>(tw, tw.act[L - 2], gcur, nullptr, gw + surr_woff(L - 1), gb + surr_boff(L - 1), s);
    dgrad<kSurrHidden, kSurrOut>(tw, gcur, nullptr, gnext, w + surr_woff(L - 1), s);
    std::swap(gcur, gnext);
    for (int j = L - 2; j >= 1; --j) {
        wgrad<kSurrHidden, kSurrHidden>(tw, tw.act[j - 1], gcur, tw.act[j], gw + surr_woff(j), gb + surr_boff(j), s);
        dgrad<kSurrHidden, kSurrHidden>(tw, gcur, tw.act[j], gnext, w + surr_woff(j), s);
        std::swap(gcur, gnext);
    }
    wgrad<kSurrIn, kSurrHidden>(tw, feat, gcur, tw.act[0], gw + surr_woff(0), gb + surr_boff(0), s);
}

}  // namespace

hipError_t surr_forward_tables(const SurrImg* imgs, const SurrTile* tiles, int ntiles, const uint16_t* w_bf16,
                               const float* bias, const float* feat, float* out, float* tmp0, float* tmp1,
                               hipStream_t s) {
    float* act[kSurrLayers];
    for (int j = 0; j < kSurrLayers - 1; ++j) act[j] = (j & 1) ? tmp1 : tmp0;
    act[kSurrLayers - 1] = out;
    forward_layers(imgs, tiles, ntiles, feat, act, w_bf16, bias, s);
    return hipGetLastError();
}

}  // namespace i360

using namespace i360;

namespace {

int Finish(i360hip_ctx* ctx, const Status& st) {
    if (!st.ok()) set_error(ctx, st.msg);
    return st.code;
}

hipStream_t StreamOf(i360hip_ctx* ctx, void* s) { return s ? static_cast<hipStream_t>(s) : ctx->stream; }

bool check_dims(int n, int h, int w, Status& st) {
    if (n <= 0 || h <= 0 || w <= 0 || static_cast<long long>(n) * h * w > (1ll << 26)) {
        st.set(I360HIP_INVALID_PARAM, "surrogate batch %d x %d x %d outside 1 .. 2^26 pixels", n, h, w);
        return false;
    }
    return true;
}

}  // namespace

int i360hip_surrogate_forward(i360hip_ctx* ctx, const uint16_t* weights, const float* biases, const float* feat,
                              int n, int h, int w, float* out, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (!weights || !biases || !feat || !out) st.set(I360HIP_INVALID_PARAM, "Null pointer argument.");
    else if (check_dims(n, h, w, st) && hipSetDevice(ctx->device) == hipSuccess) {
        hipStream_t s = StreamOf(ctx, stream);
        TrainWork tw{};
        if (train_work(ctx, n, h, w, tw, st, s)) {
            ProfScope ps(ctx, s, "surrogate_forward", 0.0);
            if (surr_forward_tables(tw.imgs, tw.tiles, tw.ntiles, weights, biases, feat, out, tw.g0, tw.g1, s) !=
                hipSuccess)
                st.set(I360HIP_KERNEL_LAUNCH_FAILED, "surrogate forward launch failed");
        }
    }
    return Finish(ctx, st);
}

int i360hip_surrogate_gradients(i360hip_ctx* ctx, const uint16_t* weights, const float* biases, const float* feat,
                                const float* target, int n, int h, int w, float* grad_weights, float* grad_biases,
                                float* loss, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (!weights || !biases || !feat || !target || !grad_weights || !grad_biases || !loss) {
        st.set(I360HIP_INVALID_PARAM, "Null pointer argument.");
        return Finish(ctx, st);
    }
    if (!check_dims(n, h, w, st) || hipSetDevice(ctx->device) != hipSuccess) return Finish(ctx, st);
    hipStream_t s = StreamOf(ctx, stream);
    TrainWork tw{};
    if (!train_work(ctx, n, h, w, tw, st, s)) return Finish(ctx, st);
    gradients(tw, weights, biases, feat, target, s);
    I360_TRY(st, hipGetLastError());
    I360_TRY(st, hipMemcpyAsync(grad_weights, tw.grad, sizeof(float) * I360HIP_SURR_WEIGHTS, hipMemcpyDeviceToDevice, s));
    I360_TRY(st, hipMemcpyAsync(grad_biases, tw.grad + I360HIP_SURR_WEIGHTS, sizeof(float) * I360HIP_SURR_BIASES,
                                hipMemcpyDeviceToDevice, s));
    I360_TRY(st, hipMemcpyAsync(loss, tw.loss, sizeof(float), hipMemcpyDeviceToDevice, s));
    return Finish(ctx, st);
}

int i360hip_surrogate_train_step(i360hip_ctx* ctx, float* master_weights, float* biases, float* adam_state,
                                 uint16_t* weights, const float* feat, const float* target, int n, int h, int w,
                                 float lr, int step, float* loss, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (!master_weights || !biases || !adam_state || !weights || !feat || !target || !loss) {
        st.set(I360HIP_INVALID_PARAM, "Null pointer argument.");
        return Finish(ctx, st);
    }
    if (step < 1 || !(lr > 0.0f)) {
        st.set(I360HIP_INVALID_PARAM, "step must be >= 1 and lr > 0");
        return Finish(ctx, st);
    }
    if (!check_dims(n, h, w, st) || hipSetDevice(ctx->device) != hipSuccess) return Finish(ctx, st);
    hipStream_t s = StreamOf(ctx, stream);
    TrainWork tw{};
    if (!train_work(ctx, n, h, w, tw, st, s)) return Finish(ctx, st);
    {
        ProfScope ps(ctx, s, "surrogate_train_step", 0.0);
        gradients(tw, weights, biases, feat, target, s);
        const float c1 = 1.0f / (1.0f - std::pow(0.9f, static_cast<float>(step)));
        const float c2 = 1.0f / (1.0f - std::pow(0.999f, static_cast<float>(step)));
        constexpr int W = I360HIP_SURR_WEIGHTS, B = I360HIP_SURR_BIASES;
        surr_adam_kernel<<<(W + 255) / 256, 256, 0, s>>>(master_weights, adam_state, adam_state + W, tw.grad,
                                                         weights, W, lr, c1, c2);
        surr_adam_kernel<<<1, 256, 0, s>>>(biases, adam_state + 2 * W, adam_state + 2 * W + B, tw.grad + W,
                                           nullptr, B, lr, c1, c2);
    }
    I360_TRY(st, hipGetLastError());
    I360_TRY(st, hipMemcpyAsync(loss, tw.loss, sizeof(float), hipMemcpyDeviceToDevice, s));
    return Finish(ctx, st);
}

int i360hip_surrogate_round_bf16(i360hip_ctx* ctx, const float* master_weights, uint16_t* weights, void* stream) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    std::lock_guard<std::mutex> lk(ctx->mu);
    Status st;
    if (!master_weights || !weights) {
        st.set(I360HIP_INVALID_PARAM, "Null pointer argument.");
        return Finish(ctx, st);
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return I360HIP_DEVICE_ERROR;
    constexpr int W = I360HIP_SURR_WEIGHTS;
    surr_to_bf16_kernel<<<(W + 255) / 256, 256, 0, StreamOf(ctx, stream)>>>(master_weights, weights, W);
    I360_TRY(st, hipGetLastError());
    return Finish(ctx, st);
}
