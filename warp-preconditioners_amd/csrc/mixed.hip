// Mixed (indefinite) harmonic systems: assembly of the interleaved 2/3-field
// block matrix with its shifted SPD companion H = K + weight^(1/order) M, and the
// per-vertex glue of the matching-Schur and shifted-block preconditioners
// (mixed.py:14-133, 325-383).  Vectors of the system are interleaved: vertex i
// holds components x[i*order + c].  Each kernel is one thread per vertex; all
// arithmetic is the reference's expression order without contraction.

#include "wpc_common.h"

namespace wpc {

constexpr int kMixedThreads = 256;

// mixed.py:19-57.  Triplet i is the vertex's diagonal block, triplet n+e the
// block of K entry e; duplicates (K's own diagonal) are summed by the caller's
// triplet assembly.
template <class T, int ORDER>
__global__ __launch_bounds__(kMixedThreads) void mixed_build_kernel(
    int n, const int* __restrict__ off, const int* __restrict__ col, const T* __restrict__ val,
    const T* __restrict__ mass, T weight, T shift, int* __restrict__ rows, int* __restrict__ cols,
    T* __restrict__ blocks, T* __restrict__ hvalues, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int BB = ORDER * ORDER;
    const T m = mass[i];
    if (m <= T(0) || !isfinite(m)) atomicAdd(bad, 1);
    rows[i] = i;
    cols[i] = i;
    T block[BB];
#pragma unroll
    for (int k = 0; k < BB; ++k) block[k] = T(0);
    block[0] = weight * m;
    if (ORDER == 2) {
        block[1 * ORDER + 1] = -m;
    } else {
        block[1 * ORDER + 2] = -m;
        block[2 * ORDER + 1] = -m;
    }
#pragma unroll
    for (int k = 0; k < BB; ++k) blocks[(long)i * BB + k] = block[k];
    hvalues[i] = shift * m;
    for (int e = off[i]; e < off[i + 1]; ++e) {
        const long t = (long)n + e;
        rows[t] = i;
        cols[t] = col[e];
        const T k = -val[e];
#pragma unroll
        for (int q = 0; q < BB; ++q) block[q] = T(0);
        if (ORDER == 2) {
            block[0 * ORDER + 1] = k;
            block[1 * ORDER + 0] = k;
        } else {
            block[0 * ORDER + 2] = k;
            block[2 * ORDER + 0] = k;
            block[1 * ORDER + 1] = k;
        }
#pragma unroll
        for (int q = 0; q < BB; ++q) blocks[t * BB + q] = block[q];
        hvalues[t] = val[e];
    }
}

// mixed.py:66-79: out = x[:, c] (or x[:, c] / mass)
template <class T>
__global__ __launch_bounds__(kMixedThreads) void component_kernel(int n, int order,
                                                                  const T* __restrict__ x, int c,
                                                                  const T* __restrict__ mass,
                                                                  T* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T v = x[(long)i * order + c];
    out[i] = mass != nullptr ? v / mass[i] : v;
}

// mixed.py:81-86, 118-121: out = mass * x, or x / mass (out may alias x)
template <class T>
__global__ __launch_bounds__(kMixedThreads) void mass_kernel(int n, const T* x, const T* __restrict__ mass,
                                                             int divide, T* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = divide ? x[i] / mass[i] : mass[i] * x[i];
}

// mixed.py:88-102: out = x[:, 0] - k1 (+ k2 for three fields)
template <class T>
__global__ __launch_bounds__(kMixedThreads) void rhs_kernel(int n, int order, const T* __restrict__ x,
                                                            const T* __restrict__ k1,
                                                            const T* __restrict__ k2,
                                                            T* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v = x[(long)i * order] - k1[i];
    if (order == 3) v += k2[i];
    out[i] = v;
}

// mixed.py:104-116: out = -(ku + x[:, c]) / mass
template <class T>
__global__ __launch_bounds__(kMixedThreads) void auxiliary_kernel(int n, int order,
                                                                  const T* __restrict__ x, int c,
                                                                  const T* __restrict__ ku,
                                                                  const T* __restrict__ mass,
                                                                  T* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = -(ku[i] + x[(long)i * order + c]) / mass[i];
}

// mixed.py:123-143: z = alpha * (u, a[, lam]) + beta * y; alpha == 0 reads none
// of the fields, beta == 0 never reads y (y and z may alias).
template <class T>
__global__ __launch_bounds__(kMixedThreads) void finish_kernel(int n, int order, const T* __restrict__ u,
                                                               const T* __restrict__ a,
                                                               const T* __restrict__ lam, const T* y,
                                                               T* z, T alpha, T beta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v[3] = {T(0), T(0), T(0)};
    if (alpha != T(0)) {
        v[0] = u[i] * alpha;
        v[1] = a[i] * alpha;
        if (order == 3) v[2] = lam[i] * alpha;
    }
    for (int c = 0; c < order; ++c) {
        const long k = (long)i * order + c;
        T w = v[c];
        if (beta != T(0)) w = w + beta * y[k];
        z[k] = w;
    }
}

// mixed.py:330-345: split an interleaved vector into the shifted-block fields
template <class T>
__global__ __launch_bounds__(kMixedThreads) void unpack_kernel(int n, int order, const T* __restrict__ x,
                                                               T* __restrict__ a, T* __restrict__ b,
                                                               T* __restrict__ c, T scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T* xi = x + (long)i * order;
    if (order == 2) {
        a[i] = (xi[0] / scale - scale * xi[1]) * T(0.5);
        b[i] = (-xi[0] / scale - scale * xi[1]) * T(0.5);
    } else {
        a[i] = -xi[2];
        b[i] = -xi[1];
        c[i] = -xi[0];
    }
}

// mixed.py:347-372
template <class T>
__global__ __launch_bounds__(kMixedThreads) void shifted_finish_kernel(int n, int order,
                                                                       const T* __restrict__ a,
                                                                       const T* __restrict__ b,
                                                                       const T* __restrict__ c,
                                                                       const T* y, T* z, T scale,
                                                                       T alpha, T beta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v[3] = {T(0), T(0), T(0)};
    if (alpha != T(0)) {
        if (order == 2) {
            v[0] = a[i] / scale * alpha;
            v[1] = b[i] * scale * alpha;
        } else {
            v[0] = a[i] * alpha;
            v[1] = b[i] * alpha;
            v[2] = c[i] * alpha;
        }
    }
    for (int q = 0; q < order; ++q) {
        const long k = (long)i * order + q;
        T w = v[q];
        if (beta != T(0)) w = w + beta * y[k];
        z[k] = w;
    }
}

// mixed.py:147-160: out[i, c] = scale[c] * x[i, c] (out may alias x)
template <class T>
__global__ __launch_bounds__(kMixedThreads) void vector_scale_kernel(long count, int order, const T* x,
                                                                     T* out, T s0, T s1, T s2) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int c = (int)(k % order);
    const T s = c == 0 ? s0 : (c == 1 ? s1 : s2);
    out[k] = x[k] * s;
}

}  // namespace wpc

using namespace wpc;

#define MIXED_TYPE_SWITCH(dt, CALL)                          \
    if ((dt) == WPC_F64) { using T = double; CALL; }         \
    else if ((dt) == WPC_F32) { using T = float; CALL; }     \
    else return WPC_EDTYPE;

#define MIXED_GRID(n) dim3(blocks_for((n), kMixedThreads)), dim3(kMixedThreads), 0, as_stream(stream)

extern "C" {

int wpc_mixed_build(int n, int order, const int* offsets, const int* columns, const void* values,
                    const void* mass, double weight, double shift, int* rows, int* cols,
                    void* blocks, void* hvalues, int* bad, int dtype, void* stream) {
    if (n <= 0) return WPC_OK;
    if (order != 2 && order != 3) return WPC_EARG;
    MIXED_TYPE_SWITCH(dtype, {
        if (order == 2)
            hipLaunchKernelGGL((mixed_build_kernel<T, 2>), MIXED_GRID(n), n, offsets, columns,
                               (const T*)values, (const T*)mass, (T)weight, (T)shift, rows, cols,
                               (T*)blocks, (T*)hvalues, bad);
        else
            hipLaunchKernelGGL((mixed_build_kernel<T, 3>), MIXED_GRID(n), n, offsets, columns,
                               (const T*)values, (const T*)mass, (T)weight, (T)shift, rows, cols,
                               (T*)blocks, (T*)hvalues, bad);
    })
    return launch_status();
}

int wpc_mixed_component(int n, int order, const void* x, int c, const void* mass, void* out,
                        int dtype, void* stream) {
    if (n <= 0) return WPC_OK;
    if (order < 1 || c < 0 || c >= order) return WPC_EARG;
    MIXED_TYPE_SWITCH(dtype, hipLaunchKernelGGL((component_kernel<T>), MIXED_GRID(n), n, order,
                                                (const T*)x, c, (const T*)mass, (T*)out))
    return launch_status();
}

int wpc_mass_multiply(int n, const void* x, const void* mass, int divide, void* out, int dtype,
                      void* stream) {
    if (n <= 0) return WPC_OK;
    MIXED_TYPE_SWITCH(dtype, hipLaunchKernelGGL((mass_kernel<T>), MIXED_GRID(n), n, (const T*)x,
                                                (const T*)mass, divide, (T*)out))
    return launch_status();
}

int wpc_mixed_rhs(int n, int order, const void* x, const void* k1, const void* k2, void* out,
                  int dtype, void* stream) {
    if (n <= 0) return WPC_OK;
    if (order != 2 && order != 3) return WPC_EARG;
    MIXED_TYPE_SWITCH(dtype, hipLaunchKernelGGL((rhs_kernel<T>), MIXED_GRID(n), n, order,
                                                (const T*)x, (const T*)k1, (const T*)k2, (T*)out))
    return launch_status();
}

int wpc_mixed_auxiliary(int n, int order, const void* x, int c, const void* ku, const void* mass,
                        void* out, int dtype, void* stream) {
    if (n <= 0) return WPC_OK;
    if ((order != 2 && order != 3) || c < 0 || c >= order) return WPC_EARG;
    MIXED_TYPE_SWITCH(dtype, hipLaunchKernelGGL((auxiliary_kernel<T>), MIXED_GRID(n), n, order,
                                                (const T*)x, c, (const T*)ku, (const T*)mass,
                                                (T*)out))
    return launch_status();
}

int wpc_mixed_finish(int n, int order, const void* u, const void* a, const void* lam,
                     const void* y, void* z, double alpha, double beta, int dtype, void* stream) {
    if (n <= 0) return WPC_OK;
    if (order != 2 && order != 3) return WPC_EARG;
    MIXED_TYPE_SWITCH(dtype, hipLaunchKernelGGL((finish_kernel<T>), MIXED_GRID(n), n, order,
                                                (const T*)u, (const T*)a, (const T*)lam,
                                                (const T*)y, (T*)z, (T)alpha, (T)beta))
    return launch_status();
}

int wpc_shifted_unpack(int n, int order, const void* x, void* a, void* b, void* c, double scale,
                       int dtype, void* stream) {
    if (n <= 0) return WPC_OK;
    if (order != 2 && order != 3) return WPC_EARG;
    MIXED_TYPE_SWITCH(dtype, hipLaunchKernelGGL((unpack_kernel<T>), MIXED_GRID(n), n, order,
                                                (const T*)x, (T*)a, (T*)b, (T*)c, (T)scale))
    return launch_status();
}

int wpc_shifted_finish(int n, int order, const void* a, const void* b, const void* c,
                       const void* y, void* z, double scale, double alpha, double beta, int dtype,
                       void* stream) {
    if (n <= 0) return WPC_OK;
    if (order != 2 && order != 3) return WPC_EARG;
    MIXED_TYPE_SWITCH(dtype, hipLaunchKernelGGL((shifted_finish_kernel<T>), MIXED_GRID(n), n, order,
                                                (const T*)a, (const T*)b, (const T*)c, (const T*)y,
                                                (T*)z, (T)scale, (T)alpha, (T)beta))
    return launch_status();
}

int wpc_vector_scale(long count, int order, const void* x, void* out, double s0, double s1,
                     double s2, int dtype, void* stream) {
    if (count <= 0) return WPC_OK;
    if (order < 1 || order > 3 || count % order != 0) return WPC_EARG;
    MIXED_TYPE_SWITCH(dtype, hipLaunchKernelGGL((vector_scale_kernel<T>), MIXED_GRID(count), count,
                                                order, (const T*)x, (T*)out, (T)s0, (T)s1, (T)s2))
    return launch_status();
}

}  // extern "C"
