// CPU stand-in of the HIP runtime, for running the library's host code
// (planner, table upload, launch sequences, stream capture and graph
// bookkeeping) on the CPU under AddressSanitizer and UBSan.  The library's
// .hip sources are compiled host-only and linked against this file instead
// of libamdhip64.
//
// Device memory is host memory: small blocks from the heap, large ones
// reserved with mmap (never touched unless the host copies into them).
// Kernels do not run: a launch checks its configuration and stream and
// counts as one operation on that stream.  Every call the library makes
// is checked, and a violation aborts with a message naming it:
//   * copies and fills must lie inside one live allocation;
//   * frees must name a live allocation;
//   * stream capture follows the runtime's rules: a stream joins a capture
//     by waiting on an event recorded in it; at the end of capture every
//     joined stream's last operation must be ordered before the origin's
//     end (tracked with vector clocks); no synchronising, allocating or
//     freeing call while a capture is active; no memset node in a captured
//     graph (fills recorded into graphs have given wrong replays on the
//     target machines, so the library must not record them);
//   * graph launches, uploads and destruction name live objects, outside
//     capture.
// At exit it prints one summary line (counts of launches, captures, graph
// launches), which the test reads.
#include <hip/hip_runtime_api.h>
#include <sys/mman.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

namespace {

[[noreturn]] void fail(const char* fmt, ...) {
    std::va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "HIP STUB VIOLATION: ");
    std::vfprintf(stderr, fmt, ap);
    std::fputc('\n', stderr);
    va_end(ap);
    std::fflush(stderr);
    std::abort();
}

constexpr size_t kMapThreshold = size_t(1) << 24;  // blocks from 16 MiB up are reserved, not touched

struct Alloc {
    size_t size;
    bool mapped;
};

struct State;
State& S();
std::map<uintptr_t, Alloc>& allocs();

const std::pair<const uintptr_t, Alloc>* find_alloc(const void* p) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    auto it = allocs().upper_bound(a);
    if (it == allocs().begin()) return nullptr;
    --it;
    if (a >= it->first + it->second.size) return nullptr;
    return &*it;
}

bool is_device(const void* p) { return find_alloc(p) != nullptr; }

void check_range(const void* p, size_t n, const char* what) {
    if (n == 0) return;
    const auto* al = find_alloc(p);
    if (!al) fail("%s: %p is not in a live device allocation", what, p);
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    if (a + n > al->first + al->second.size)
        fail("%s: [%p, +%zu) runs %zu bytes past its allocation of %zu bytes", what, p, n,
             static_cast<size_t>(a + n - (al->first + al->second.size)), al->second.size);
}

// vector clock: stream -> last operation of that stream this one has observed
typedef std::map<hipStream_t, uint64_t> Clock;

void merge(Clock& into, const Clock& from) {
    for (const auto& kv : from) {
        uint64_t& v = into[kv.first];
        if (kv.second > v) v = kv.second;
    }
}

struct Capture {
    hipStream_t origin;
    std::set<hipStream_t> members;
    size_t nodes = 0;
    bool invalid = false;
};

struct StreamState {
    int capture = 0;  // capture id, 0: not capturing
    uint64_t ops = 0;
    uint64_t ops_at_join = 0;  // ops when it joined its capture
    Clock clock;
};

struct EventState {
    int capture = 0;
    Clock clock;  // what a waiter observes
    bool recorded = false;
};

// All state lives in one object created on first use and never destroyed:
// the compiler-emitted module constructors register kernels and device
// variables before this file's globals would be constructed, and the
// summary at exit still reads it.
struct State {
    std::recursive_mutex mu;
    std::map<uintptr_t, Alloc> allocs;
    std::map<hipStream_t, StreamState> streams;  // created streams (and the null stream)
    std::map<hipEvent_t, EventState> events;
    std::map<int, Capture> captures;
    int next_capture = 1;
    std::set<hipGraph_t> graphs;
    std::set<hipGraphExec_t> execs;
    std::map<const void*, size_t> vars;    // registered device variables: host shadow -> size
    std::map<const void*, void*> var_dev;  // their device copies
    int device = 0;
    uint64_t launches = 0, graph_launches = 0, captures_done = 0, copies = 0;
};
State& S() {
    static State* s = new State;
    return *s;
}
thread_local std::vector<std::tuple<dim3, dim3, size_t, hipStream_t>> t_config;
std::map<uintptr_t, Alloc>& allocs() { return S().allocs; }

StreamState& stream_of(hipStream_t s, const char* what) {
    auto it = S().streams.find(s);
    if (it == S().streams.end()) {
        if (s == nullptr) return S().streams[nullptr];
        fail("%s: stream %p was not created (or was destroyed)", what, static_cast<void*>(s));
    }
    return it->second;
}

bool any_capture() {
    for (const auto& kv : S().captures)
        if (!kv.second.invalid) return true;
    return false;
}

void no_capture(const char* what) {
    if (any_capture()) fail("%s called while a stream capture is active", what);
}

// One operation on stream s (a node when s is capturing).
void op(hipStream_t s, const char* what) {
    StreamState& st = stream_of(s, what);
    ++st.ops;
    st.clock[s] = st.ops;
    if (st.capture) ++S().captures[st.capture].nodes;
}

void* dev_alloc(size_t n) {
    const size_t sz = n ? n : 1;
    void* p = nullptr;
    bool mapped = false;
    if (sz >= kMapThreshold) {
        p = mmap(nullptr, sz, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (p == MAP_FAILED) return nullptr;
        mapped = true;
    } else {
        p = std::calloc(1, sz);
        if (!p) return nullptr;
    }
    S().allocs[reinterpret_cast<uintptr_t>(p)] = Alloc{sz, mapped};
    return p;
}

struct Summary {
    ~Summary() {
        std::printf("HIP STUB OK: %llu launches, %llu captures, %llu graph launches, %llu copies, %zu live allocations\n",
                    static_cast<unsigned long long>(S().launches), static_cast<unsigned long long>(S().captures_done),
                    static_cast<unsigned long long>(S().graph_launches), static_cast<unsigned long long>(S().copies),
                    S().allocs.size());
        std::fflush(stdout);
    }
} g_summary;

}  // namespace

extern "C" {

// ---- registration and launch (emitted by the compiler for <<<>>>) ----------
void** __hipRegisterFatBinary(const void*) {
    static void* handle = nullptr;
    return &handle;
}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, void*, void*, void*, void*, int*) {}
void __hipRegisterVar(void**, void* var, char*, char*, int, size_t size, int, int) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    S().vars[var] = size;
}

hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shmem, hipStream_t stream) {
    t_config.emplace_back(grid, block, shmem, stream);
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* grid, dim3* block, size_t* shmem, hipStream_t* stream) {
    if (t_config.empty()) fail("__hipPopCallConfiguration without a pushed configuration");
    std::tie(*grid, *block, *shmem, *stream) = t_config.back();
    t_config.pop_back();
    return hipSuccess;
}

hipError_t hipLaunchKernel(const void* fn, dim3 grid, dim3 block, void** args, size_t shmem, hipStream_t stream) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (!fn) fail("hipLaunchKernel: null kernel");
    if (grid.x == 0 || grid.y == 0 || grid.z == 0 || grid.x > 0x7fffffffu || grid.y > 65535u || grid.z > 65535u)
        fail("hipLaunchKernel: grid %u x %u x %u out of range", grid.x, grid.y, grid.z);
    const unsigned long long threads = 1ull * block.x * block.y * block.z;
    if (threads == 0 || threads > 1024 || block.x > 1024 || block.y > 1024 || block.z > 1024)
        fail("hipLaunchKernel: block %u x %u x %u out of range", block.x, block.y, block.z);
    if (shmem > 160 * 1024) fail("hipLaunchKernel: %zu bytes of dynamic LDS", shmem);
    if (!args) fail("hipLaunchKernel: null argument array");
    op(stream, "hipLaunchKernel");
    ++S().launches;
    return hipSuccess;
}

// ---- device ---------------------------------------------------------------
hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int d) {
    if (d != 0) return hipErrorInvalidDevice;
    S().device = d;
    return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
    *d = S().device;
    return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600* p, int d) {
    if (d != 0) return hipErrorInvalidDevice;
    std::memset(p, 0, sizeof(*p));
    std::snprintf(p->name, sizeof(p->name), "CPU stand-in of a gfx950 device");
    std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950:sramecc+:xnack-");
    p->totalGlobalMem = size_t(288) << 30;
    p->sharedMemPerBlock = 160 * 1024;
    p->maxSharedMemoryPerMultiProcessor = 160 * 1024;
    p->regsPerBlock = 512 * 64;
    p->warpSize = 64;
    p->maxThreadsPerBlock = 1024;
    p->maxThreadsDim[0] = p->maxThreadsDim[1] = p->maxThreadsDim[2] = 1024;
    p->maxGridSize[0] = 0x7fffffff;
    p->maxGridSize[1] = p->maxGridSize[2] = 65535;
    p->clockRate = 2400000;
    p->multiProcessorCount = 256;
    p->maxThreadsPerMultiProcessor = 2048;
    p->l2CacheSize = 4 << 20;
    p->major = 9;
    p->minor = 5;
    p->concurrentKernels = 1;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t attr, int d) {
    if (d != 0) return hipErrorInvalidDevice;
    switch (attr) {
        case hipDeviceAttributeMultiprocessorCount: *v = 256; break;
        case hipDeviceAttributeMaxThreadsPerBlock: *v = 1024; break;
        case hipDeviceAttributeWarpSize: *v = 64; break;
        case hipDeviceAttributeMaxSharedMemoryPerBlock: *v = 160 * 1024; break;
        default: *v = 0; break;
    }
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* freeb, size_t* total) {
    *total = size_t(288) << 30;
    *freeb = size_t(280) << 30;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    no_capture("hipDeviceSynchronize");
    return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "hipSuccess" : "hip stub error"; }
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void* f, int block, size_t) {
    if (!f || block <= 0 || block > 1024) fail("hipOccupancyMaxActiveBlocksPerMultiprocessor: block %d", block);
    *n = 2048 / block < 1 ? 1 : (2048 / block > 8 ? 8 : 2048 / block);
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** dev, const void* symbol) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    auto it = S().vars.find(symbol);
    if (it == S().vars.end()) fail("hipGetSymbolAddress: %p is not a registered device variable", symbol);
    void*& d = S().var_dev[symbol];
    if (!d) d = dev_alloc(it->second);
    *dev = d;
    return hipSuccess;
}

// ---- memory ----------------------------------------------------------------
hipError_t hipMalloc(void** p, size_t n) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    no_capture("hipMalloc");
    if (!p) fail("hipMalloc: null output");
    *p = dev_alloc(n);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    no_capture("hipFree");
    if (!p) return hipSuccess;
    auto it = S().allocs.find(reinterpret_cast<uintptr_t>(p));
    if (it == S().allocs.end()) fail("hipFree: %p is not the start of a live allocation", p);
    if (it->second.mapped) munmap(p, it->second.size);
    else std::free(p);
    S().allocs.erase(it);
    return hipSuccess;
}

static void copy_checked(void* dst, const void* src, size_t n, hipMemcpyKind kind, const char* what) {
    const bool dd = is_device(dst), sd = is_device(src);
    switch (kind) {
        case hipMemcpyHostToDevice:
            if (!dd) fail("%s host-to-device: destination %p is not device memory", what, dst);
            break;
        case hipMemcpyDeviceToHost:
            if (!sd) fail("%s device-to-host: source %p is not device memory", what, src);
            break;
        case hipMemcpyDeviceToDevice:
            if (!dd || !sd) fail("%s device-to-device: %p <- %p not both device memory", what, dst, src);
            break;
        default:
            break;
    }
    if (dd) check_range(dst, n, what);
    if (sd) check_range(src, n, what);
    if (n) std::memmove(dst, src, n);  // host sides are checked by AddressSanitizer
    ++S().copies;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind kind) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    no_capture("hipMemcpy");
    copy_checked(dst, src, n, kind, "hipMemcpy");
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind kind, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    copy_checked(dst, src, n, kind, "hipMemcpyAsync");
    op(s, "hipMemcpyAsync");
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int v, size_t n, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (stream_of(s, "hipMemsetAsync").capture) fail("hipMemsetAsync recorded into a captured graph");
    check_range(dst, n, "hipMemsetAsync");
    const auto* al = find_alloc(dst);
    if (!al->second.mapped) std::memset(dst, v, n);  // reserved blocks stay untouched
    op(s, "hipMemsetAsync");
    return hipSuccess;
}

// ---- streams and events ----------------------------------------------------------------
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    *s = reinterpret_cast<hipStream_t>(new char[1]);
    S().streams[*s];
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    StreamState& st = stream_of(s, "hipStreamDestroy");
    if (st.capture) fail("hipStreamDestroy of a capturing stream");
    S().streams.erase(s);
    delete[] reinterpret_cast<char*>(s);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (stream_of(s, "hipStreamSynchronize").capture) fail("hipStreamSynchronize of a capturing stream");
    no_capture("hipStreamSynchronize");
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    *e = reinterpret_cast<hipEvent_t>(new char[1]);
    S().events[*e];
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (!S().events.erase(e)) fail("hipEventDestroy: %p is not a live event", static_cast<void*>(e));
    delete[] reinterpret_cast<char*>(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    auto it = S().events.find(e);
    if (it == S().events.end()) fail("hipEventRecord: %p is not a live event", static_cast<void*>(e));
    StreamState& st = stream_of(s, "hipEventRecord");
    it->second.capture = st.capture;
    it->second.clock = st.clock;
    it->second.recorded = true;
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    auto it = S().events.find(e);
    if (it == S().events.end()) fail("hipStreamWaitEvent: %p is not a live event", static_cast<void*>(e));
    StreamState& st = stream_of(s, "hipStreamWaitEvent");
    const EventState& ev = it->second;
    if (ev.capture) {
        Capture& c = S().captures[ev.capture];
        if (c.invalid) fail("hipStreamWaitEvent on an event of an ended capture");
        if (st.capture == 0) {  // fork: the stream joins the capture
            st.capture = ev.capture;
            st.ops_at_join = st.ops;
            c.members.insert(s);
        } else if (st.capture != ev.capture) {
            fail("hipStreamWaitEvent joins two captures");
        }
    } else if (st.capture) {
        // a capturing stream waiting on an event recorded outside capture:
        // the graph cannot express it
        if (ev.recorded) fail("a capturing stream waits on an event recorded outside the capture");
    }
    merge(st.clock, ev.clock);
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    auto it = S().events.find(e);
    if (it == S().events.end()) fail("hipEventSynchronize: %p is not a live event", static_cast<void*>(e));
    if (it->second.capture) fail("hipEventSynchronize on an event recorded in a capture");
    no_capture("hipEventSynchronize");
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (!S().events.count(a) || !S().events.count(b)) fail("hipEventElapsedTime on a dead event");
    if (S().events[a].capture || S().events[b].capture) fail("hipEventElapsedTime on events recorded in a capture");
    *ms = 1.0f;
    return hipSuccess;
}

// ---- capture and graphs ----------------------------------------------------------------
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (s == nullptr) fail("hipStreamBeginCapture on the null stream");
    StreamState& st = stream_of(s, "hipStreamBeginCapture");
    if (st.capture) fail("hipStreamBeginCapture on a capturing stream");
    const int id = S().next_capture++;
    Capture& c = S().captures[id];
    c.origin = s;
    c.members.insert(s);
    st.capture = id;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    StreamState& st = stream_of(s, "hipStreamEndCapture");
    if (!st.capture) fail("hipStreamEndCapture on a stream that is not capturing");
    Capture& c = S().captures[st.capture];
    if (c.origin != s) fail("hipStreamEndCapture on a stream that did not begin the capture");
    for (hipStream_t m : c.members) {
        if (m == s) continue;
        const uint64_t last = S().streams[m].ops;
        if (last == S().streams[m].ops_at_join) continue;  // joined, captured nothing
        auto it = st.clock.find(m);
        if (it == st.clock.end() || it->second < last)
            fail("hipStreamEndCapture: stream %p's last captured operation is not joined into the origin",
                 static_cast<void*>(m));
    }
    for (hipStream_t m : c.members) S().streams[m].capture = 0;
    c.invalid = true;
    ++S().captures_done;
    *g = reinterpret_cast<hipGraph_t>(new char[1]);
    S().graphs.insert(*g);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (!S().graphs.erase(g)) fail("hipGraphDestroy: %p is not a live graph", static_cast<void*>(g));
    delete[] reinterpret_cast<char*>(g);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (!S().graphs.count(g)) fail("hipGraphInstantiate: %p is not a live graph", static_cast<void*>(g));
    *x = reinterpret_cast<hipGraphExec_t>(new char[1]);
    S().execs.insert(*x);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (!S().execs.erase(x)) fail("hipGraphExecDestroy: %p is not a live executable graph", static_cast<void*>(x));
    delete[] reinterpret_cast<char*>(x);
    return hipSuccess;
}
hipError_t hipGraphUpload(hipGraphExec_t x, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (!S().execs.count(x)) fail("hipGraphUpload: %p is not a live executable graph", static_cast<void*>(x));
    if (stream_of(s, "hipGraphUpload").capture) fail("hipGraphUpload into a capture");
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
    std::lock_guard<std::recursive_mutex> lk(S().mu);
    if (!S().execs.count(x)) fail("hipGraphLaunch: %p is not a live executable graph", static_cast<void*>(x));
    if (stream_of(s, "hipGraphLaunch").capture) fail("hipGraphLaunch into a capture");
    op(s, "hipGraphLaunch");
    ++S().graph_launches;
    return hipSuccess;
}

}  // extern "C"
