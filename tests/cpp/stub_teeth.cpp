// Shows that the CPU stand-in of the HIP runtime (hip_runtime_stub.cpp)
// catches the mistakes it checks for: each argument commits one of them
// ("ok" commits none: a fork and join under capture, as the library's
// chunk groups do).  Run by the host-render sanitizer test, which expects
// every mistake to abort with its message.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const char* what = argv[1];
    hipStream_t a = nullptr, b = nullptr;
    hipEvent_t e = nullptr, j = nullptr;
    void* p = nullptr;
    void* q = nullptr;
    char host[256] = {};
    hipStreamCreateWithFlags(&a, hipStreamNonBlocking);
    hipStreamCreateWithFlags(&b, hipStreamNonBlocking);
    hipEventCreate(&e);
    hipEventCreate(&j);
    hipMalloc(&p, 64);
    hipMalloc(&q, 64);
    hipGraph_t g = nullptr;
    if (!std::strcmp(what, "ok") || !std::strcmp(what, "unjoined")) {
        hipStreamBeginCapture(a, hipStreamCaptureModeThreadLocal);
        hipEventRecord(e, a);
        hipStreamWaitEvent(b, e, 0);  // fork
        hipMemcpyAsync(q, p, 64, hipMemcpyDeviceToDevice, b);
        if (!std::strcmp(what, "ok")) {
            hipEventRecord(j, b);
            hipStreamWaitEvent(a, j, 0);  // join
        }
        hipStreamEndCapture(a, &g);
        hipGraphExec_t x = nullptr;
        hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        hipGraphUpload(x, a);
        hipGraphLaunch(x, a);
        hipStreamSynchronize(a);
        hipGraphExecDestroy(x);
        hipGraphDestroy(g);
    } else if (!std::strcmp(what, "memset_in_capture")) {
        hipStreamBeginCapture(a, hipStreamCaptureModeThreadLocal);
        hipMemsetAsync(p, 0, 64, a);
    } else if (!std::strcmp(what, "copy_overrun")) {
        hipMemcpy(p, host, 65, hipMemcpyHostToDevice);
    } else if (!std::strcmp(what, "free_inside")) {
        hipFree(static_cast<char*>(p) + 8);
    } else if (!std::strcmp(what, "sync_in_capture")) {
        hipStreamBeginCapture(a, hipStreamCaptureModeThreadLocal);
        hipStreamSynchronize(b);
    } else {
        return 2;
    }
    hipFree(p);
    hipFree(q);
    hipEventDestroy(e);
    hipEventDestroy(j);
    hipStreamDestroy(a);
    hipStreamDestroy(b);
    std::printf("teeth %s: no violation\n", what);
    return 0;
}
