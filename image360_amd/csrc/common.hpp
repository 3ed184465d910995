// Internal shared definitions of the HIP backend (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "i360hip.h"

namespace i360 {

// A device buffer owned by a context, grown on demand and reused.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Launch-time status of one operation: the first failure wins.
struct Status {
    int code = I360HIP_SUCCESS;
    char msg[512] = {};
    bool ok() const { return code == I360HIP_SUCCESS; }
    void set(int c, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
};

inline const char* HipName(hipError_t e) { return hipGetErrorString(e); }

// Maps a HIP error onto the reference's result codes.
inline int CodeOf(hipError_t e) {
    switch (e) {
        case hipSuccess: return I360HIP_SUCCESS;
        case hipErrorOutOfMemory: return I360HIP_OUT_OF_MEMORY;
        case hipErrorInvalidValue:
        case hipErrorInvalidDevicePointer: return I360HIP_INVALID_PARAM;
        case hipErrorLaunchFailure:
        case hipErrorInvalidConfiguration:
        case hipErrorInvalidDeviceFunction: return I360HIP_KERNEL_LAUNCH_FAILED;
        case hipErrorNoDevice:
        case hipErrorInvalidDevice: return I360HIP_NO_GPU;
        default: return I360HIP_DEVICE_ERROR;
    }
}

#define I360_TRY(st, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            (st).set(::i360::CodeOf(e_), "%s failed: %s (%s:%d)", #expr, ::i360::HipName(e_), \
                     __FILE__, __LINE__);                                                    \
            return (st).code;                                                                \
        }                                                                                    \
    } while (0)

// Grid size for `n` items at `per_block` items per block, capped by the
// 2^31-1 x-dimension limit.
inline unsigned GridFor(size_t n, size_t per_block) {
    size_t g = (n + per_block - 1) / per_block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

}  // namespace i360

namespace i360 {
// Per-launch timing records (HIP events on the launch stream), kept only
// while profiling is enabled on the context.
struct ProfRecord {
    const char* name;
    hipEvent_t start, stop;
    double alg_bytes;
};
}  // namespace i360

struct i360hip_ctx {
    void* render_cache = nullptr;  // cached render plan / graph (render_float.hip)
    bool profiling = false;
    std::vector<i360::ProfRecord> prof;
    std::vector<hipEvent_t> event_pool;
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    char last_error[512] = {};
    hipDeviceProp_t props{};
    std::vector<i360::DevBuf> bufs;  // indexed scratch slots
    // learned surrogate parameters for the SURROGATE render form
    const uint16_t* surr_weights = nullptr;  // device copy (a scratch slot)
    const float* surr_biases = nullptr;
    float surr_threshold = 1e-5f;
    float surr_max_threshold = 0.0f;  // per-pixel bound of the fallback test (0: off)
    int surr_tiles = 0, surr_fallback = -1;  // last surrogate render
    // launch configuration of the batched render (i360hip_render_autotune;
    // 0 = built-in default)
    int tune_groups = 0, tune_strip = 0, tune_share = 0;
    // export composite tables last uploaded (export.hip): argument hash,
    // the device block they live in, and the offsets of the square offsets
    // and chunk lists in it
    uint64_t export_key = 0;
    void* export_dev = nullptr;
    size_t export_off[2] = {0, 0};
    std::vector<i360::DevBuf> pinned;
};

namespace i360 {
// Global-address-space views of device pointers.  Pointers loaded from
// tables are generic; accessing through them emits flat instructions,
// which count against the LDS wait counter too and so turn every LDS wait
// into a wait for all outstanding memory traffic.
#define I360_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ I360_GLOBAL T* gptr(T* p) {
    return (I360_GLOBAL T*)(p);
}
template <typename T>
__device__ __forceinline__ const I360_GLOBAL T* gptr(const T* p) {
    return (const I360_GLOBAL T*)(p);
}

// Workgroup barrier that orders LDS only: waits for this wave's LDS
// operations (lgkmcnt), not for its outstanding global loads and stores,
// so prefetches and fire-and-forget stores stay in flight across it.  Use
// only where the data exchanged between threads lives in LDS.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Context-owned device buffer `slot` of at least `bytes` bytes (grown on
// demand, kept until the context is destroyed).
void* ctx_scratch(i360hip_ctx* ctx, int slot, size_t bytes, Status& st);
// Records the text of a failing call: in the context (i360hip_last_error,
// shared by every thread that calls it) and for the calling thread
// (i360hip_thread_error, which no other thread's failure can replace).
void set_error(i360hip_ctx* ctx, const char* msg);
// Looks up, once per device, the device address of the allocation-free
// statistics entry point's scratch array (color.hip); `device` is current.
bool color_scratch_setup(int device);
// Drops the context's cached render plan and graph.
void render_release(i360hip_ctx* ctx);
}  // namespace i360

namespace i360 {
// Brackets one kernel launch with events when profiling is on.
struct ProfScope {
    i360hip_ctx* ctx;
    hipStream_t s;
    ProfRecord rec{};
    bool on;
    ProfScope(i360hip_ctx* c, hipStream_t st, const char* name, double bytes);
    ~ProfScope();
};
}  // namespace i360
