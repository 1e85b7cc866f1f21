// Factored squared-Laplacian operator (L M^-1 L + w M)[f, f] as two sparse passes.
//
// Pass 1 streams every row of L once (values, columns, gathered x) and writes
// tmp = M^-1 (L x~); pass 2 streams the free rows of L again and gathers tmp.
// Both passes are HBM-bound; algorithmic bytes for one product with nvec vectors
// of xb bytes and vb-byte values: 2*nnz*(vb+4) + 8*(n+1) + (2 + 5)*n*nvec*xb.

#include "wpc_common.h"

namespace wpc {

__global__ __launch_bounds__(256) void index_free_kernel(int n_free, const int* free, int n,
                                                         int* index, int* status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_free) return;
    const int j = free[i];
    if (j < 0 || j >= n) {
        atomicAdd(status, 1);
    } else {
        const int old = atomicCAS(index + j, -1, i);
        if (old != -1) atomicAdd(status + 1, 1);
    }
}

template <class T>
__global__ __launch_bounds__(256) void mass_inverse_kernel(int n, const T* mass, T* inverse,
                                                           int* status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T m = mass[i];
    if (m > T(0) && isfinite(m)) inverse[i] = T(1) / m;
    else atomicAdd(status + 2, 1);
}

// squared_laplacian.py:37-65 (forward).
template <class T, int NV>
__global__ __launch_bounds__(256) void forward_kernel(int n_full, const int* __restrict__ off,
                                                      const int* __restrict__ col,
                                                      const T* __restrict__ val,
                                                      const T* __restrict__ inverse,
                                                      const int* __restrict__ index,
                                                      const T* __restrict__ x, T* __restrict__ tmp,
                                                      int boundary, int lanes_log2) {
    const int L = 1 << lanes_log2;
    const long row = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> lanes_log2;
    const int sub = threadIdx.x & (L - 1);
    T acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = T(0);
    const bool live = row < n_full;
    if (live) {
        for (int e = off[row] + sub; e < off[row + 1]; e += L) {
            const int c0 = col[e];
            const int c = index ? index[c0] : c0;
            long src = -1;
            if (boundary) {
                if (c < 0) src = c0;
            } else {
                if (c >= 0) src = c;
            }
            if (src >= 0) {
                const T a = val[e];
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] += a * x[src * NV + v];
            }
        }
    }
    for (int m = L >> 1; m > 0; m >>= 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += shfl_xor(acc[v], m, kWave);
    }
    if (live && sub == 0) {
        const T inv = inverse[row];
#pragma unroll
        for (int v = 0; v < NV; ++v) tmp[row * NV + v] = inv * acc[v];
    }
}

// squared_laplacian.py:67-98 (backward).
template <class T, int NV>
__global__ __launch_bounds__(256) void backward_kernel(
    int n_free, const int* __restrict__ off, const int* __restrict__ col, const T* __restrict__ val,
    const int* __restrict__ free, const T* __restrict__ mass, const T* x, T mass_weight,
    const T* __restrict__ tmp, const T* y, T* z, T alpha, T beta, int lanes_log2) {
    const int L = 1 << lanes_log2;
    const long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> lanes_log2;
    const int sub = threadIdx.x & (L - 1);
    T acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = T(0);
    const bool live = i < n_free;
    const long row = live ? (free ? free[i] : i) : 0;
    if (live && alpha != T(0)) {
        for (int e = off[row] + sub; e < off[row + 1]; e += L) {
            const T a = val[e];
            const long c = col[e];
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += a * tmp[c * NV + v];
        }
    }
    for (int m = L >> 1; m > 0; m >>= 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += shfl_xor(acc[v], m, kWave);
    }
    if (live && sub == 0) {
        const T mw = (mass_weight != T(0) && alpha != T(0)) ? mass_weight * mass[row] : T(0);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            T r = acc[v];
            if (mass_weight != T(0) && alpha != T(0)) r += mw * x[i * NV + v];
            r *= alpha;
            if (beta != T(0)) r += beta * y[i * NV + v];
            z[i * NV + v] = r;
        }
    }
}

static int lg_of(int lanes) {
    int lg = 0;
    while ((1 << lg) < lanes) ++lg;
    return (1 << lg) == lanes && lg <= 6 ? lg : -1;
}

template <class T, int NV>
static int sqlap_launch(int n_full, const int* off, const int* col, const void* val,
                        const void* inverse, const int* index, int n_free, const int* free,
                        const void* mass, double mw, void* tmp, const void* x, const void* y,
                        void* z, double alpha, double beta, int boundary, int lanes,
                        hipStream_t s) {
    const int lg = lg_of(lanes);
    if (lg < 0) return WPC_EARG;
    if (alpha != 0.0 && n_full > 0)
        hipLaunchKernelGGL((forward_kernel<T, NV>), dim3(blocks_for((long)n_full << lg, 256)),
                           dim3(256), 0, s, n_full, off, col, (const T*)val, (const T*)inverse,
                           index, (const T*)x, (T*)tmp, boundary, lg);
    if (n_free > 0)
        hipLaunchKernelGGL((backward_kernel<T, NV>), dim3(blocks_for((long)n_free << lg, 256)),
                           dim3(256), 0, s, n_free, off, col, (const T*)val, free, (const T*)mass,
                           (const T*)x, (T)(boundary ? 0.0 : mw), (const T*)tmp, (const T*)y,
                           (T*)z, (T)alpha, (T)beta, lg);
    return launch_status();
}

}  // namespace wpc

using namespace wpc;

extern "C" {

int wpc_sqlap_index_free(int n_free, const int* free_indices, int n, int* index, int* status,
                         void* stream) {
    if (n_free <= 0) return WPC_OK;
    hipLaunchKernelGGL(index_free_kernel, dim3(blocks_for(n_free, 256)), dim3(256), 0,
                       as_stream(stream), n_free, free_indices, n, index, status);
    return launch_status();
}

int wpc_sqlap_mass_inverse(int n, const void* mass, void* inverse, int dtype, int* status,
                           void* stream) {
    if (n <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(mass_inverse_kernel<double>, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                           n, (const double*)mass, (double*)inverse, status);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(mass_inverse_kernel<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                           n, (const float*)mass, (float*)inverse, status);
    else
        return WPC_EDTYPE;
    return launch_status();
}

int wpc_sqlap_apply(int n_full, const int* offsets, const int* columns, const void* values,
                    int dtype, const void* inverse, const int* index, int n_free,
                    const int* free_indices, const void* mass, double mass_weight, void* tmp,
                    const void* x, const void* y, void* z, int nvec, double alpha, double beta,
                    int boundary, int lanes, void* stream) {
    hipStream_t s = as_stream(stream);
    if (lanes == 0) lanes = 8;
#define SQ(T, NV) return sqlap_launch<T, NV>(n_full, offsets, columns, values, inverse, index, n_free, free_indices, mass, mass_weight, tmp, x, y, z, alpha, beta, boundary, lanes, s)
#define SQNV(T)                      \
    switch (nvec) {                  \
        case 1: SQ(T, 1);            \
        case 2: SQ(T, 2);            \
        case 3: SQ(T, 3);            \
        case 4: SQ(T, 4);            \
        case 5: SQ(T, 5);            \
        case 6: SQ(T, 6);            \
        case 7: SQ(T, 7);            \
        case 8: SQ(T, 8);            \
        default: return WPC_EARG;    \
    }
    if (dtype == WPC_F64) SQNV(double)
    if (dtype == WPC_F32) SQNV(float)
#undef SQNV
#undef SQ
    return WPC_EDTYPE;
}

}  // extern "C"
