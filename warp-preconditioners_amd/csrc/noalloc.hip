// The allocation-free variant of wpc_fsai_apply: the two-pass product's
// temporary G x lives in a file-scope device array instead of a caller buffer,
// for callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  Same kernels as
// wpc_fsai_apply (spmv.hip), so the results are identical bit for bit.
//
// Each GPU the library runs on holds its own copy of `scratch`
// (WPC_NOALLOC_SCRATCH_BYTES) until the process ends.  Its device address is
// looked up once per device, with that device current, on the first call there;
// host code never passes the array's name to a kernel (in host code the name is
// a stand-in, not a device address).  Calls that overlap on one GPU share the
// array and race: the caller serialises them (one stream, or events).

#include <mutex>

#include "wpc_common.h"
#include "wpc.h"

namespace wpc {

int csr_spmv(int n, const int* off, const int* col, const void* val, int vdt, const float* scale,
             const void* x, const void* y, void* z, int xdt, int nvec, double alpha, double beta,
             int lanes, hipStream_t s);

namespace {

static __device__ __align__(256) unsigned char scratch[WPC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_scratch[kMaxDevices] = {};

// The current device's address of `scratch` (looked up once per device), or
// NULL with *rc set to the HIP error.
void* device_scratch(int* rc) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *rc = (int)e;
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        *rc = WPC_EARG;
        return nullptr;
    }
    std::lock_guard<std::mutex> hold(g_lock);
    if (g_scratch[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            *rc = e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
            return nullptr;
        }
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

long vector_bytes(int dtype) {
    switch (dtype) {
        case WPC_F64: return 8;
        case WPC_F32: return 4;
        case WPC_F16: case WPC_BF16: return 2;
    }
    return 0;
}

}  // namespace
}  // namespace wpc

using namespace wpc;

extern "C" int wpc_fsai_apply_noalloc(int n, const int* g_offsets, const int* g_columns,
                                      const void* g_values, const int* gt_offsets,
                                      const int* gt_columns, const void* gt_values,
                                      int factor_dtype, const float* row_scale2, const void* x,
                                      const void* y, void* z, int vec_dtype, int nvec,
                                      double alpha, double beta, int lanes, void* stream) {
    const long vb = vector_bytes(vec_dtype);
    if (vb == 0) return WPC_EDTYPE;
    if (n < 0 || nvec < 1 || nvec > 8) return WPC_EARG;
    if ((long)n * nvec * vb > (long)WPC_NOALLOC_SCRATCH_BYTES) return WPC_ESCRATCH;
    int rc = WPC_OK;
    void* tmp = device_scratch(&rc);
    if (tmp == nullptr) return rc;
    hipStream_t s = as_stream(stream);
    if (lanes == 0) lanes = 8;
    // The first pass writes every element of tmp the second pass reads (alpha == 0
    // reads neither), so the array's earlier contents never matter.
    if (alpha != 0.0) {
        rc = csr_spmv(n, g_offsets, g_columns, g_values, factor_dtype, row_scale2, x, nullptr,
                      tmp, vec_dtype, nvec, 1.0, 0.0, lanes, s);
        if (rc) return rc;
    }
    return csr_spmv(n, gt_offsets, gt_columns, gt_values, factor_dtype, nullptr, tmp, y, z,
                    vec_dtype, nvec, alpha, beta, lanes, s);
}
