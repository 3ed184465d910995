// Event-bracketed kernel timing (i360hip_profile_enable / _read).

#include <cstring>
#include <map>
#include <string>

#include "common.hpp"

namespace i360 {

namespace {
hipEvent_t TakeEvent(i360hip_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

ProfScope::ProfScope(i360hip_ctx* c, hipStream_t st, const char* name, double bytes)
    : ctx(c), s(st), on(c && c->profiling) {
    if (!on) return;
    rec.name = name;
    rec.alg_bytes = bytes;
    rec.start = TakeEvent(ctx);
    rec.stop = TakeEvent(ctx);
    if (!rec.start || !rec.stop) { on = false; return; }
    (void)hipEventRecord(rec.start, s);
}

ProfScope::~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(rec.stop, s);
    ctx->prof.push_back(rec);
}

}  // namespace i360

using namespace i360;

extern "C" {

int i360hip_profile_enable(i360hip_ctx* ctx, int enable) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    for (auto& r : ctx->prof) {
        (void)hipEventSynchronize(r.stop);
        ctx->event_pool.push_back(r.start);
        ctx->event_pool.push_back(r.stop);
    }
    ctx->prof.clear();
    ctx->profiling = enable != 0;
    return I360HIP_SUCCESS;
}

int i360hip_profile_read(i360hip_ctx* ctx, i360hip_kernel_stat* stats, int max_stats, int* num_stats) {
    if (!ctx) return I360HIP_DEVICE_ERROR;
    if (!num_stats || (max_stats > 0 && !stats)) return I360HIP_INVALID_PARAM;
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    std::map<std::string, i360hip_kernel_stat> agg;
    for (auto& r : ctx->prof) {
        if (hipEventSynchronize(r.stop) != hipSuccess) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) continue;
        auto& a = agg[r.name];
        std::snprintf(a.name, sizeof(a.name), "%s", r.name);
        a.launches += 1;
        a.total_ms += ms;
        a.alg_bytes += r.alg_bytes;
    }
    int k = 0;
    for (auto& kv : agg) {
        if (k < max_stats) stats[k] = kv.second;
        ++k;
    }
    *num_stats = k < max_stats ? k : max_stats;
    return I360HIP_SUCCESS;
}

}  // extern "C"
