// Drives the library's render entry points over the CPU stand-in of the
// HIP runtime (hip_runtime_stub.cpp), so the planner, table uploads,
// launch sequences, chunk-group forks and joins, stream capture and graph
// bookkeeping run on the CPU under the sanitizers the test builds with.
// Kernels do not run; outputs are not checked here (the GPU tests do).
//
// Usage: render_host_driver GEOMETRY_FILE
// The geometry file (written by the test): "frames chunks bands", then one
// line per frame "w h gain h0 .. h8", then one line per chunk
// "x y w h count c0 c1 ...".
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "i360hip.h"

namespace {

i360hip_ctx* g_ctx = nullptr;

void check(int rc, const char* what) {
    if (rc != I360HIP_SUCCESS) {
        std::fprintf(stderr, "FAILED %s: %d (%s)\n", what, rc, g_ctx ? i360hip_last_error(g_ctx) : "");
        std::exit(1);
    }
}

void* dev(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, n) != hipSuccess || !p) {
        std::fprintf(stderr, "hipMalloc(%zu) failed\n", n);
        std::exit(1);
    }
    return p;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s GEOMETRY_FILE\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "r");
    if (!f) return 2;
    int nf = 0, nc = 0, bands = 0;
    if (std::fscanf(f, "%d %d %d", &nf, &nc, &bands) != 3 || nf <= 0 || nc <= 0) return 2;
    std::vector<i360hip_frame> frames(nf);
    for (auto& fr : frames) {
        std::memset(&fr, 0, sizeof(fr));
        if (std::fscanf(f, "%d %d %f", &fr.width, &fr.height, &fr.gain) != 3) return 2;
        for (double& v : fr.h)
            if (std::fscanf(f, "%lf", &v) != 1) return 2;
        fr.format = I360HIP_PIX_F16;
        fr.pixels = dev(static_cast<size_t>(fr.width) * fr.height * 6);
    }
    std::vector<i360hip_chunk> chunks(nc);
    std::vector<int> contrib;
    size_t out_values = 0, owner_values = 0;
    for (auto& c : chunks) {
        std::memset(&c, 0, sizeof(c));
        if (std::fscanf(f, "%d %d %d %d %d", &c.x, &c.y, &c.width, &c.height, &c.count) != 5) return 2;
        c.first = static_cast<int>(contrib.size());
        for (int k = 0; k < c.count; ++k) {
            int v = 0;
            if (std::fscanf(f, "%d", &v) != 1) return 2;
            contrib.push_back(v);
        }
        out_values += static_cast<size_t>(c.width) * c.height * 3;
        owner_values += static_cast<size_t>(c.width) * c.height;
    }
    std::fclose(f);
    auto* out = static_cast<float*>(dev(out_values * 4));
    auto* owner = static_cast<int32_t*>(dev(owner_values * 4));
    auto set_outputs = [&](bool owners) {
        size_t o = 0, w = 0;
        for (auto& c : chunks) {
            c.out = out + o;
            c.owner = owners ? owner + w : nullptr;
            o += static_cast<size_t>(c.width) * c.height * 3;
            w += static_cast<size_t>(c.width) * c.height;
        }
    };
    check(i360hip_create(0, &g_ctx), "create");
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return 1;

    auto render = [&](int mode, int count, const char* what) {
        std::printf("render %s (mode %d, %d chunks)\n", what, mode, count);
        std::fflush(stdout);
        check(i360hip_render_chunks(g_ctx, frames.data(), nf, chunks.data(), count, contrib.data(), bands, mode, s),
              what);
        if (hipStreamSynchronize(s) != hipSuccess) std::exit(1);
    };
    auto tune = [&](int groups, int strip, int share) {
        i360hip_tuning t{groups, strip, share, 0.0};
        check(i360hip_render_set_tuning(g_ctx, &t), "set_tuning");
    };

    // batched form: capture, then replays; RGB48 output; staggered and
    // side-by-side chunk groups; one group; owners
    set_outputs(false);
    tune(4, 4, 1);
    setenv("I360HIP_RENDER_STAGGER", "1", 1);
    for (int r = 0; r < 3; ++r) render(I360HIP_RENDER_BATCHED, nc, "batched, four staggered groups");
    render(I360HIP_RENDER_BATCHED | I360HIP_RENDER_OUT_RGB48, nc, "batched RGB48");
    setenv("I360HIP_RENDER_STAGGER", "0", 1);
    tune(3, 4, 2);
    for (int r = 0; r < 2; ++r) render(I360HIP_RENDER_BATCHED, nc, "batched, three groups side by side");
    setenv("I360HIP_RENDER_STAGGER", "1", 1);
    tune(1, 8, 1);
    render(I360HIP_RENDER_BATCHED, nc, "batched, one group, strip 8");
    tune(8, 0, 1);
    set_outputs(true);
    render(I360HIP_RENDER_BATCHED, nc, "batched, eight groups, owners");
    // per-chunk forms on a prefix of the chunks
    const int few = nc < 3 ? nc : 3;
    render(I360HIP_RENDER_PRECISION, nc, "precision");
    render(I360HIP_RENDER_PLAIN, few, "plain");
    render(I360HIP_RENDER_FP64, few < 2 ? few : 2, "fp64");
    // profiling brackets every launch with events
    check(i360hip_profile_enable(g_ctx, 1), "profile_enable");
    tune(4, 4, 1);
    render(I360HIP_RENDER_BATCHED, nc, "batched, profiled");
    check(i360hip_profile_enable(g_ctx, 0), "profile_disable");
    // surrogate form (parameters of zeros: what matters is the launch
    // sequence), with and without the per-pixel bound
    set_outputs(false);
    void* w = dev(2 * static_cast<size_t>(I360HIP_SURR_WEIGHTS));
    void* b = dev(4 * static_cast<size_t>(I360HIP_SURR_BIASES));
    check(i360hip_surrogate_set(g_ctx, static_cast<const uint16_t*>(w), static_cast<const float*>(b), 1e-5f),
          "surrogate_set");
    render(I360HIP_RENDER_SURROGATE, nc, "surrogate");
    check(i360hip_surrogate_set_max_error(g_ctx, 1e-6f), "surrogate_set_max_error");
    render(I360HIP_RENDER_SURROGATE | I360HIP_RENDER_OUT_RGB48, nc, "surrogate RGB48, bounded");
    int tiles = 0, fb = 0;
    check(i360hip_surrogate_stats(g_ctx, &tiles, &fb), "surrogate_stats");
    // the autotuner over a prefix (every candidate is a capture and replays)
    i360hip_tuning best{};
    check(i360hip_render_autotune(g_ctx, frames.data(), nf, chunks.data(), few, contrib.data(), bands, &best),
          "autotune");
    std::printf("autotune: groups %d strip %d share %d\n", best.chunk_groups, best.chamfer_strip, best.grid_share);
    render(I360HIP_RENDER_BATCHED, few, "batched after autotune");

    if (hipStreamDestroy(s) != hipSuccess) return 1;
    i360hip_destroy(g_ctx);
    g_ctx = nullptr;
    (void)hipFree(w);
    (void)hipFree(b);
    (void)hipFree(out);
    (void)hipFree(owner);
    for (auto& fr : frames) (void)hipFree(const_cast<void*>(fr.pixels));
    std::printf("driver done\n");
    return 0;
}
