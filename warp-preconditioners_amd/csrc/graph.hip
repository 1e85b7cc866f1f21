// Stream capture of launch sequences into HIP graphs, for callers that drive the
// library from a scripting host: capturing on the caller's stream needs no device
// synchronization (unlike a framework graph context that drains the device and its
// allocator first), so a serving loop can capture one request's chunk of solver
// steps while other requests keep running on other streams.  The library's
// launches allocate nothing, so every captured node refers to caller-owned memory.

#include "wpc_common.h"

using wpc::as_stream;

extern "C" {

int wpc_graph_begin(void* stream) {
    hipError_t e = hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) (void)hipGetLastError();  // leave no error for the next launch check
    return e == hipSuccess ? WPC_OK : (int)e;
}

// A failed capture (e.g. invalidated) ends here with no graph, the stream out of
// capture mode and the thread's error state cleared, so the caller can launch
// the same work directly.
int wpc_graph_end(void* stream, void** graph_exec) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(as_stream(stream), &g);
    if (e != hipSuccess) {
        if (g != nullptr) hipGraphDestroy(g);
        (void)hipGetLastError();
        return (int)e;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return (int)e;
    *graph_exec = (void*)x;
    return WPC_OK;
}

int wpc_graph_launch(void* graph_exec, void* stream) {
    hipError_t e = hipGraphLaunch((hipGraphExec_t)graph_exec, as_stream(stream));
    return e == hipSuccess ? WPC_OK : (int)e;
}

// A stream of the caller's own (non-blocking, default priority).  Frameworks that
// hand out streams from a shared pool can give two owners the same stream; a
// capture must run on a stream that no other work uses.
int wpc_stream_create(void** stream) {
    if (stream == nullptr) return WPC_EARG;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return (int)e;
    }
    *stream = (void*)s;
    return WPC_OK;
}

int wpc_stream_destroy(void* stream) {
    if (stream == nullptr) return WPC_OK;
    hipError_t e = hipStreamDestroy(as_stream(stream));
    return e == hipSuccess ? WPC_OK : (int)e;
}

int wpc_graph_destroy(void* graph_exec) {
    if (graph_exec == nullptr) return WPC_OK;
    hipError_t e = hipGraphExecDestroy((hipGraphExec_t)graph_exec);
    return e == hipSuccess ? WPC_OK : (int)e;
}

}  // extern "C"
