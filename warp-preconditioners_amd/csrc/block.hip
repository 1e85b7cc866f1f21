// Block Jacobi and synchronous block ILU(0) for b x b BSR (b = 1..4).
//
// Blocks are row-major b*b values; vectors are b-interleaved.  Setup kernels run
// one thread per stored block (sweep) or per block row; applies are one thread per
// block row and HBM-bound (the L/U sweeps stream the factor once per step).

#include "wpc_common.h"

namespace wpc {

template <class T, int B>
__device__ T det_b(const T* a) {
    if (B == 1) return a[0];
    if (B == 2) return a[0] * a[3] - a[1] * a[2];
    if (B == 3)
        return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
               a[2] * (a[3] * a[7] - a[4] * a[6]);
    T total = T(0);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        T m[9];
        int k = 0;
#pragma unroll
        for (int r = 1; r < 4; ++r)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
                if (cc != c) m[k++] = a[r * 4 + cc];
        const T d = det_b<T, 3>(m);
        total += (c % 2 ? -a[c] : a[c]) * d;
    }
    return total;
}

// Adjugate / determinant.
template <class T, int B>
__device__ void inverse_b(const T* a, T* out) {
    if (B == 1) {
        out[0] = T(1) / a[0];
        return;
    }
    const T d = det_b<T, B>(a);
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            T m[9];
            int k = 0;
#pragma unroll
            for (int rr = 0; rr < B; ++rr)
#pragma unroll
                for (int cc = 0; cc < B; ++cc)
                    if (rr != r && cc != c) m[k++] = a[rr * B + cc];
            T cof = B == 2 ? m[0] : (B == 3 ? det_b<T, 2>(m) : det_b<T, 3>(m));
            if ((r + c) % 2) cof = -cof;
            out[c * B + r] = cof / d;
        }
}

// block_ilu.py:16-43 (safe_inverse): max-abs scaling, determinant test, inverse.
template <class T, int B>
__device__ void safe_inverse(const T* a, T tolerance, T* inverse, int* bad) {
    T scale = T(0);
#pragma unroll
    for (int k = 0; k < B * B; ++k) scale = fmax(scale, fabs(a[k]));
#pragma unroll
    for (int k = 0; k < B * B; ++k) inverse[k] = T(0);
    int fails = 0;
    if (scale > T(0) && isfinite(scale)) {
        T nrm[B * B];
#pragma unroll
        for (int k = 0; k < B * B; ++k) nrm[k] = a[k] / scale;
        const T det = det_b<T, B>(nrm);
        if (isfinite(det) && fabs(det) > tolerance) {
            if (B == 1) {
                inverse[0] = T(1) / a[0];
            } else {
                inverse_b<T, B>(nrm, inverse);
#pragma unroll
                for (int k = 0; k < B * B; ++k) inverse[k] = inverse[k] / scale;
            }
        } else {
            fails += 1;
        }
    } else {
        fails += 1;
    }
#pragma unroll
    for (int k = 0; k < B * B; ++k)
        if (!isfinite(inverse[k])) fails += 1;
    if (fails) atomicAdd(bad, fails);
}

template <class T, int B>
__device__ __forceinline__ void matmul_b(const T* a, const T* b, T* out) {
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            T s = T(0);
#pragma unroll
            for (int k = 0; k < B; ++k) s += a[r * B + k] * b[k * B + c];
            out[r * B + c] = s;
        }
}

// block_ilu.py:45-61 (topology).
__global__ __launch_bounds__(256) void topology_kernel(int nb, const int* off, const int* col,
                                                       int* rows, int* diagonal, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    int found = -1;
    for (int e = off[i]; e < off[i + 1]; ++e) {
        rows[e] = i;
        if (col[e] == i) found = e;
    }
    diagonal[i] = found;
    if (found < 0) atomicAdd(bad, 1);
}

// block_ilu.py:63-73 (invert).
template <class T, int B>
__global__ __launch_bounds__(256) void invert_kernel(int nb, const T* values, const int* diagonal,
                                                     T* inverse, T tolerance, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb || diagonal[i] < 0) return;
    T a[B * B], inv[B * B];
#pragma unroll
    for (int k = 0; k < B * B; ++k) a[k] = values[(long)diagonal[i] * B * B + k];
    safe_inverse<T, B>(a, tolerance, inv, bad);
#pragma unroll
    for (int k = 0; k < B * B; ++k) inverse[(long)i * B * B + k] = inv[k];
}

// block_ilu.py:75-104 (diagonal_only, compact storage).
template <class T, int B>
__global__ __launch_bounds__(256) void diagonal_only_kernel(int nb, const int* off, const int* col,
                                                            const T* values, T* inverse,
                                                            T tolerance, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    int lo = off[i], hi = off[i + 1];
    const int end = hi;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < i) lo = mid + 1;
        else hi = mid;
    }
    if (lo < end && col[lo] == i) {
        T a[B * B], inv[B * B];
#pragma unroll
        for (int k = 0; k < B * B; ++k) a[k] = values[(long)lo * B * B + k];
        safe_inverse<T, B>(a, tolerance, inv, bad);
#pragma unroll
        for (int k = 0; k < B * B; ++k) inverse[(long)i * B * B + k] = inv[k];
    } else {
        atomicAdd(bad, 1);
    }
}

// block_ilu.py:106-118 (initialize).
template <class T, int B>
__global__ __launch_bounds__(256) void ilu_initialize_kernel(long nnzb, const int* rows,
                                                             const int* col, const T* values,
                                                             const T* inverse, T* factors) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnzb) return;
    T a[B * B], out[B * B];
#pragma unroll
    for (int k = 0; k < B * B; ++k) a[k] = values[e * B * B + k];
    if (col[e] < rows[e]) {
        matmul_b<T, B>(a, inverse + (long)col[e] * B * B, out);
    } else {
#pragma unroll
        for (int k = 0; k < B * B; ++k) out[k] = a[k];
    }
#pragma unroll
    for (int k = 0; k < B * B; ++k) factors[e * B * B + k] = out[k];
}

// block_ilu.py:120-150 (sweep).
template <class T, int B>
__global__ __launch_bounds__(256) void ilu_sweep_kernel(long nnzb, const int* __restrict__ off,
                                                        const int* __restrict__ col,
                                                        const int* __restrict__ rows,
                                                        const T* __restrict__ values,
                                                        const T* __restrict__ old,
                                                        const T* __restrict__ inverse,
                                                        T* __restrict__ out, T relaxation) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnzb) return;
    const int i = rows[e], j = col[e];
    const int kmax = i < j ? i : j;
    T value[B * B];
#pragma unroll
    for (int k = 0; k < B * B; ++k) value[k] = values[e * B * B + k];
    for (int ik = off[i]; ik < off[i + 1]; ++ik) {
        const int k = col[ik];
        if (k >= kmax) break;
        int lo = off[k], hi = off[k + 1];
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (col[mid] < j) lo = mid + 1;
            else hi = mid;
        }
        if (lo < off[k + 1] && col[lo] == j) {
            T prod[B * B];
            matmul_b<T, B>(old + (long)ik * B * B, old + (long)lo * B * B, prod);
#pragma unroll
            for (int q = 0; q < B * B; ++q) value[q] -= prod[q];
        }
    }
    if (j < i) {
        T tmp[B * B];
        matmul_b<T, B>(value, inverse + (long)j * B * B, tmp);
#pragma unroll
        for (int q = 0; q < B * B; ++q) value[q] = tmp[q];
    }
#pragma unroll
    for (int q = 0; q < B * B; ++q)
        out[e * B * B + q] = (T(1) - relaxation) * old[e * B * B + q] + relaxation * value[q];
}

// block_ilu.py:152-171 (triangular).
template <class T, int B>
__global__ __launch_bounds__(256) void ilu_triangular_kernel(int nb, const int* __restrict__ off,
                                                             const int* __restrict__ col,
                                                             const T* __restrict__ factors,
                                                             const T* __restrict__ inverse,
                                                             const T* rhs, const T* old, T* out,
                                                             int lower) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    T value[B];
#pragma unroll
    for (int k = 0; k < B; ++k) value[k] = rhs[(long)i * B + k];
    for (int e = off[i]; e < off[i + 1]; ++e) {
        const int j = col[e];
        if ((lower && j < i) || (!lower && j > i)) {
            const T* f = factors + (long)e * B * B;
#pragma unroll
            for (int r = 0; r < B; ++r) {
                T s = T(0);
#pragma unroll
                for (int c = 0; c < B; ++c) s += f[r * B + c] * old[(long)j * B + c];
                value[r] -= s;
            }
        }
    }
    if (!lower) {
        const T* d = inverse + (long)i * B * B;
        T tmp[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            T s = T(0);
#pragma unroll
            for (int c = 0; c < B; ++c) s += d[r * B + c] * value[c];
            tmp[r] = s;
        }
#pragma unroll
        for (int r = 0; r < B; ++r) value[r] = tmp[r];
    }
#pragma unroll
    for (int k = 0; k < B; ++k) out[(long)i * B + k] = value[k];
}

// block_ilu.py:187-202 (diagonal_apply).
template <class T, int B>
__global__ __launch_bounds__(256) void diagonal_apply_kernel(int nb, const T* __restrict__ inverse,
                                                             const T* x, const T* y, T* z, T alpha,
                                                             T beta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    T value[B];
#pragma unroll
    for (int k = 0; k < B; ++k) value[k] = T(0);
    if (alpha != T(0)) {
        const T* d = inverse + (long)i * B * B;
        T xv[B];
#pragma unroll
        for (int c = 0; c < B; ++c) xv[c] = x[(long)i * B + c];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            T s = T(0);
#pragma unroll
            for (int c = 0; c < B; ++c) s += d[r * B + c] * xv[c];
            value[r] = alpha * s;
        }
    }
    if (beta != T(0)) {
#pragma unroll
        for (int k = 0; k < B; ++k) value[k] += beta * y[(long)i * B + k];
    }
#pragma unroll
    for (int k = 0; k < B; ++k) z[(long)i * B + k] = value[k];
}

template <class T>
__global__ __launch_bounds__(256) void count_nonfinite_kernel(long count, const T* v, int* bad,
                                                              int positive) {
    int local = 0;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (long)gridDim.x * blockDim.x) {
        const T a = v[k];
        if (!isfinite(a) || (positive && !(a > T(0)))) ++local;
    }
    if (local) atomicAdd(bad, local);
}

inline unsigned stride_grid(long count) {
    long b = (count + 255) / 256;
    if (b > 4096) b = 4096;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace wpc

using namespace wpc;

#define BLOCK_SWITCH(b, CALL)          \
    switch (b) {                       \
        case 1: { constexpr int B = 1; CALL; break; } \
        case 2: { constexpr int B = 2; CALL; break; } \
        case 3: { constexpr int B = 3; CALL; break; } \
        case 4: { constexpr int B = 4; CALL; break; } \
        default: return WPC_EARG;      \
    }

#define TYPE_SWITCH(dt, CALL)                                   \
    if ((dt) == WPC_F64) { using T = double; CALL; }            \
    else if ((dt) == WPC_F32) { using T = float; CALL; }        \
    else return WPC_EDTYPE;

extern "C" {

int wpc_block_topology(int nb, const int* offsets, const int* columns, int* rows, int* diagonal,
                       int* bad, void* stream) {
    if (nb <= 0) return WPC_OK;
    hipLaunchKernelGGL(topology_kernel, dim3(blocks_for(nb, 256)), dim3(256), 0, as_stream(stream),
                       nb, offsets, columns, rows, diagonal, bad);
    return launch_status();
}

int wpc_block_invert(int nb, int b, const void* values, const int* diagonal, void* inverse,
                     double tolerance, int* bad, int dtype, void* stream) {
    if (nb <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    TYPE_SWITCH(dtype, BLOCK_SWITCH(b, hipLaunchKernelGGL((invert_kernel<T, B>), dim3(blocks_for(nb, 256)), dim3(256), 0, s, nb, (const T*)values, diagonal, (T*)inverse, (T)tolerance, bad)))
    return launch_status();
}

int wpc_block_diagonal_only(int nb, int b, const int* offsets, const int* columns,
                            const void* values, void* inverse, double tolerance, int* bad,
                            int dtype, void* stream) {
    if (nb <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    TYPE_SWITCH(dtype, BLOCK_SWITCH(b, hipLaunchKernelGGL((diagonal_only_kernel<T, B>), dim3(blocks_for(nb, 256)), dim3(256), 0, s, nb, offsets, columns, (const T*)values, (T*)inverse, (T)tolerance, bad)))
    return launch_status();
}

int wpc_ilu_initialize(long nnzb, int b, const int* rows, const int* columns, const void* values,
                       const void* inverse, void* factors, int dtype, void* stream) {
    if (nnzb <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    TYPE_SWITCH(dtype, BLOCK_SWITCH(b, hipLaunchKernelGGL((ilu_initialize_kernel<T, B>), dim3(blocks_for(nnzb, 256)), dim3(256), 0, s, nnzb, rows, columns, (const T*)values, (const T*)inverse, (T*)factors)))
    return launch_status();
}

int wpc_ilu_sweep(long nnzb, int b, const int* offsets, const int* columns, const int* rows,
                  const void* values, const void* old, const void* inverse, void* out,
                  double relaxation, int dtype, void* stream) {
    if (nnzb <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    TYPE_SWITCH(dtype, BLOCK_SWITCH(b, hipLaunchKernelGGL((ilu_sweep_kernel<T, B>), dim3(blocks_for(nnzb, 256)), dim3(256), 0, s, nnzb, offsets, columns, rows, (const T*)values, (const T*)old, (const T*)inverse, (T*)out, (T)relaxation)))
    return launch_status();
}

int wpc_ilu_triangular(int nb, int b, const int* offsets, const int* columns, const void* factors,
                       const void* inverse, const void* rhs, const void* old, void* out,
                       int lower, int dtype, void* stream) {
    if (nb <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    TYPE_SWITCH(dtype, BLOCK_SWITCH(b, hipLaunchKernelGGL((ilu_triangular_kernel<T, B>), dim3(blocks_for(nb, 256)), dim3(256), 0, s, nb, offsets, columns, (const T*)factors, (const T*)inverse, (const T*)rhs, (const T*)old, (T*)out, lower)))
    return launch_status();
}

int wpc_block_diagonal_apply(int nb, int b, const void* inverse, const void* x, const void* y,
                             void* z, double alpha, double beta, int dtype, void* stream) {
    if (nb <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    TYPE_SWITCH(dtype, BLOCK_SWITCH(b, hipLaunchKernelGGL((diagonal_apply_kernel<T, B>), dim3(blocks_for(nb, 256)), dim3(256), 0, s, nb, (const T*)inverse, (const T*)x, (const T*)y, (T*)z, (T)alpha, (T)beta)))
    return launch_status();
}

int wpc_count_nonfinite(long count, const void* values, int dtype, int* bad, void* stream) {
    if (count <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    TYPE_SWITCH(dtype, hipLaunchKernelGGL((count_nonfinite_kernel<T>), dim3(stride_grid(count)), dim3(256), 0, s, count, (const T*)values, bad, 0))
    return launch_status();
}

int wpc_count_nonpositive(long count, const void* values, int dtype, int* bad, void* stream) {
    if (count <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    TYPE_SWITCH(dtype, hipLaunchKernelGGL((count_nonfinite_kernel<T>), dim3(stride_grid(count)), dim3(256), 0, s, count, (const T*)values, bad, 1))
    return launch_status();
}

}  // extern "C"
