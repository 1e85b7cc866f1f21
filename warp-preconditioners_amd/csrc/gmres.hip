// Krylov-basis kernels of restarted GMRES: the Gram products h = V[0..k) w and the
// basis combination w += alpha * V[0..k)^T h, for a row-major basis V (row j at
// V + j * ld).  Both are one pass over the k basis rows; the Gram product is a
// deterministic two-stage reduction (per-workgroup partials, then a fixed-order
// fold), so a solve repeats bit for bit.

#include "reduce.h"

namespace wpc {

constexpr int kGramThreads = kStreamRows;
constexpr int kGramMaxBlocks = 512;
constexpr int kGramMaxRows = 128;

inline int gram_blocks(long n) {
    long b = (n + kGramThreads * 4 - 1) / (kGramThreads * 4);
    if (b > kGramMaxBlocks) b = kGramMaxBlocks;
    return (int)(b < 1 ? 1 : b);
}

template <class T>
__global__ __launch_bounds__(kGramThreads) void gram_partials_kernel(long n, int k, const T* __restrict__ V,
                                                                     long ld, const T* __restrict__ w,
                                                                     double* __restrict__ partials) {
    __shared__ double red[kGramThreads];
    const long stride = (long)gridDim.x * kGramThreads;
    for (int j = 0; j < k; ++j) {
        const T* vj = V + (long)j * ld;
        double s = 0.0;
        for (long i = (long)blockIdx.x * kGramThreads + threadIdx.x; i < n; i += stride)
            s += (double)vj[i] * (double)w[i];
        double v[1] = {s};
        block_sum<1>(v, red);
        if (threadIdx.x == 0) partials[(long)blockIdx.x * k + j] = v[0];
    }
}

// One workgroup per basis row: h[j] = sum over blocks in a fixed order.
__global__ __launch_bounds__(kStreamRows) void gram_fold_kernel(int P, int k,
                                                                const double* __restrict__ partials,
                                                                double* __restrict__ h) {
    __shared__ double red[kStreamRows];
    const int j = blockIdx.x;
    const double s = fold_block(partials, P, k, j, red);
    if (threadIdx.x == 0) h[j] = s;
}

template <class T>
__global__ __launch_bounds__(kGramThreads) void combine_kernel(long n, int k, const T* __restrict__ V,
                                                               long ld, const double* __restrict__ h,
                                                               double alpha, T* __restrict__ w) {
    const long i = (long)blockIdx.x * kGramThreads + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += h[j] * (double)V[(long)j * ld + i];
    w[i] = (T)((double)w[i] + alpha * s);
}

}  // namespace wpc

using namespace wpc;

extern "C" {

int wpc_gram_partials(long n) { return gram_blocks(n); }

int wpc_gram(long n, int k, const void* V, long ld, const void* w, int dtype, double* partials,
             double* h, void* stream) {
    if (k <= 0) return WPC_OK;
    if (k > kGramMaxRows || n < 0 || ld < n) return WPC_EARG;
    hipStream_t s = as_stream(stream);
    const int P = gram_blocks(n);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(gram_partials_kernel<double>, dim3(P), dim3(kGramThreads), 0, s, n, k,
                           (const double*)V, ld, (const double*)w, partials);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(gram_partials_kernel<float>, dim3(P), dim3(kGramThreads), 0, s, n, k,
                           (const float*)V, ld, (const float*)w, partials);
    else
        return WPC_EDTYPE;
    hipLaunchKernelGGL(gram_fold_kernel, dim3(k), dim3(kStreamRows), 0, s, P, k, partials, h);
    return launch_status();
}

int wpc_basis_combine(long n, int k, const void* V, long ld, const double* h, double alpha,
                      void* w, int dtype, void* stream) {
    if (k <= 0 || n <= 0) return WPC_OK;
    if (k > kGramMaxRows || ld < n) return WPC_EARG;
    hipStream_t s = as_stream(stream);
    const unsigned grid = blocks_for(n, kGramThreads);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(combine_kernel<double>, dim3(grid), dim3(kGramThreads), 0, s, n, k,
                           (const double*)V, ld, h, alpha, (double*)w);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(combine_kernel<float>, dim3(grid), dim3(kGramThreads), 0, s, n, k,
                           (const float*)V, ld, h, alpha, (float*)w);
    else
        return WPC_EDTYPE;
    return launch_status();
}

}  // extern "C"
