// Accurate true residuals for certified stopping, and the vector steps of
// mixed-precision iterative refinement.
//
// Residuals are accumulated with the compensated dot product Dot2 (Ogita, Rump and
// Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput. 26(6), 2005): each
// product is split exactly with an FMA (TwoProduct) and each addition with
// Knuth's TwoSum, so a row sum is as accurate as if accumulated in twice the
// working precision and then rounded.  For the factored operator
//     A = alpha*M + K diag(inv_m) K
// the intermediate t = inv_m .* (K x) is kept as an unevaluated double-double
// (t_hi + t_lo), so r = b - alpha*M x - K t carries no cancellation noise from
// either product: for smooth fields the plain float64 evaluation of K inv_m K x is
// only ~1e-8 accurate at 1/h = 95, far above the certificates the solver needs.
//
// Both passes also return, per right-hand side, ||r||^2_{M^-1} = sum r_i^2 / m_i
// and ||x||^2_M = sum m_i x_i^2 (block partials folded by one workgroup in a fixed
// order).  Since lambda_min(M^-1 A) >= alpha for A = alpha*M + (PSD), the relative
// mass-weighted field error of x is bounded by ||r||_{M^-1} / (alpha ||x||_M).
//
// Vectors hold nvec right-hand sides interleaved [row][nvec], float64.

#include "reduce.h"

namespace wpc {

constexpr int kResidualThreads = 256;

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double z = s - a;
    e = (a - (s - z)) + (b - z);
}

__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    e = fma(a, b, -p);
}

// Dot2 accumulator: the sum is p + s, with s gathering the exact rounding errors.
struct Dot2 {
    double p, s;
    __device__ void init(double v) {
        p = v;
        s = 0.0;
    }
    __device__ void add_prod(double a, double b) {
        double h, r, t, q;
        two_prod(a, b, h, r);
        two_sum(p, h, t, q);
        p = t;
        s += q + r;
    }
    __device__ double value() const { return p + s; }
    __device__ void split(double& hi, double& lo) const { two_sum(p, s, hi, lo); }
};

// t = inv_m .* (K x) as double-double, every row and right-hand side.
template <int NV>
__global__ __launch_bounds__(kResidualThreads) void dd_forward_kernel(
    int n, const int* __restrict__ off, const int* __restrict__ col, const double* __restrict__ val,
    const double* __restrict__ inv_m, const double* __restrict__ x, double* __restrict__ t_hi,
    double* __restrict__ t_lo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Dot2 acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v].init(0.0);
    for (int e = off[i]; e < off[i + 1]; ++e) {
        const double a = val[e];
        const long c = (long)col[e] * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v].add_prod(a, x[c + v]);
    }
    const double w = inv_m[i];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double h, l, ph, pe;
        acc[v].split(h, l);
        two_prod(h, w, ph, pe);
        pe += l * w;
        double th, tl;
        two_sum(ph, pe, th, tl);
        t_hi[(long)i * NV + v] = th;
        t_lo[(long)i * NV + v] = tl;
    }
}

// r = b - diag_weight .* x - A y, with y = x (assembled operator, t_hi == nullptr)
// or y = t_hi + t_lo (factored operator, diag_weight = alpha*m).  Partials per block:
// [sum r^2/m (NV), sum m x^2 (NV)].
template <int NV>
__global__ __launch_bounds__(kResidualThreads) void dd_residual_kernel(
    int n, const int* __restrict__ off, const int* __restrict__ col, const double* __restrict__ val,
    const double* __restrict__ mass, double alpha, const double* __restrict__ x,
    const double* __restrict__ t_hi, const double* __restrict__ t_lo, const double* b, double* r,
    double* __restrict__ partials) {
    // b and r may alias (r overwrites b in place): each thread reads b only at its
    // own row before writing r there.
    __shared__ double red[kResidualThreads];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double q[2 * NV];
#pragma unroll
    for (int k = 0; k < 2 * NV; ++k) q[k] = 0.0;
    if (i < n) {
        Dot2 acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v].init(b[(long)i * NV + v]);
        const double m = mass[i];
        if (t_hi != nullptr) {
            double am, ame;
            two_prod(alpha, m, am, ame);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double xv = x[(long)i * NV + v];
                acc[v].add_prod(-am, xv);
                acc[v].add_prod(-ame, xv);
            }
            for (int e = off[i]; e < off[i + 1]; ++e) {
                const double a = -val[e];
                const long c = (long)col[e] * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    acc[v].add_prod(a, t_hi[c + v]);
                    acc[v].add_prod(a, t_lo[c + v]);
                }
            }
        } else {
            for (int e = off[i]; e < off[i + 1]; ++e) {
                const double a = -val[e];
                const long c = (long)col[e] * NV;
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v].add_prod(a, x[c + v]);
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double rv = acc[v].value();
            const double xv = x[(long)i * NV + v];
            r[(long)i * NV + v] = rv;
            q[v] = rv * rv / m;
            q[NV + v] = m * xv * xv;
        }
    }
    block_sum<2 * NV>(q, red);
    write_partials<2 * NV>(q, partials);
}

// out[k] = sum over blocks of partials[block][k], k < nq (<= 16), fixed order.
__global__ __launch_bounds__(1024) void dd_fold_kernel(int P, int nq,
                                                       const double* __restrict__ partials,
                                                       double* __restrict__ out) {
    __shared__ double red[1024 / kWave][16];
    double acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < P; b += 1024)
        for (int k = 0; k < nq; ++k) acc[k] += partials[(long)b * nq + k];
    for (int k = 0; k < nq; ++k) {
        const double s = wave_sum(acc[k]);
        if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < nq) {
        double t = 0.0;
        for (int w = 0; w < 1024 / kWave; ++w) t += red[w][threadIdx.x];
        out[threadIdx.x] = t;
    }
}

// Scale of right-hand side v for refinement: ||r_v||_{M^-1}, or 1 when zero.
__device__ __forceinline__ double refine_scale(const double* norms, int v) {
    const double s = sqrt(norms[v]);
    return s > 0.0 ? s : 1.0;
}

// out = active_v ? r / scale_v : 0 (stored as TO).
template <class TO>
__global__ __launch_bounds__(256) void refine_prepare_kernel(long total, int nvec,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ norms,
                                                             int mask, TO* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int v = (int)(i % nvec);
    const bool active = (mask >> v) & 1;
    out[i] = active ? (TO)(r[i] / refine_scale(norms, v)) : (TO)0;
}

// d = active_v ? scale_v * c : 0; x += d; optionally c_out = d (float64).
template <class TC>
__global__ __launch_bounds__(256) void refine_update_kernel(long total, int nvec,
                                                            double* __restrict__ x,
                                                            const TC* __restrict__ c,
                                                            const double* __restrict__ norms,
                                                            int mask, double* __restrict__ c_out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int v = (int)(i % nvec);
    const double d = ((mask >> v) & 1) ? refine_scale(norms, v) * (double)c[i] : 0.0;
    x[i] += d;
    if (c_out != nullptr) c_out[i] = d;
}

// values[diag(i)] += alpha * d[i]; rows without a stored diagonal count in status.
__global__ __launch_bounds__(256) void add_diagonal_kernel(int n, const int* __restrict__ off,
                                                           const int* __restrict__ col,
                                                           double* __restrict__ val,
                                                           const double* __restrict__ d,
                                                           double alpha, int* status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = off[i], hi = off[i + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < i) lo = mid + 1;
        else hi = mid;
    }
    if (lo < off[i + 1] && col[lo] == i) val[lo] += alpha * d[i];
    else atomicAdd(status, 1);
}

template <int NV>
static int dd_launch(int n, const int* off, const int* col, const double* val, const double* mass,
                     const double* inv_m, double alpha, const double* x, const double* b,
                     double* t_hi, double* t_lo, double* r, double* partials, double* norms,
                     hipStream_t s) {
    const int grid = blocks_for(n, kResidualThreads);
    if (t_hi != nullptr)
        hipLaunchKernelGGL(dd_forward_kernel<NV>, dim3(grid), dim3(kResidualThreads), 0, s, n, off,
                           col, val, inv_m, x, t_hi, t_lo);
    hipLaunchKernelGGL(dd_residual_kernel<NV>, dim3(grid), dim3(kResidualThreads), 0, s, n, off,
                       col, val, mass, alpha, x, t_hi, t_lo, b, r, partials);
    hipLaunchKernelGGL(dd_fold_kernel, dim3(1), dim3(1024), 0, s, grid, 2 * NV, partials, norms);
    return launch_status();
}

}  // namespace wpc

using namespace wpc;

extern "C" {

int wpc_residual_partials(int n) { return n > 0 ? blocks_for(n, kResidualThreads) : 1; }

int wpc_dd_residual(int n, const int* offsets, const int* columns, const double* values,
                    const double* mass, const double* inverse_mass, double alpha, const double* x,
                    const double* b, double* t_hi, double* t_lo, double* r, int nvec,
                    double* partials, double* norms, void* stream) {
    if (n <= 0) return WPC_OK;
    if ((t_hi == nullptr) != (t_lo == nullptr) || (t_hi != nullptr && inverse_mass == nullptr))
        return WPC_EARG;
    hipStream_t s = as_stream(stream);
#define DD(NV) return dd_launch<NV>(n, offsets, columns, values, mass, inverse_mass, alpha, x, b, t_hi, t_lo, r, partials, norms, s)
    switch (nvec) {
        case 1: DD(1);
        case 2: DD(2);
        case 3: DD(3);
        case 4: DD(4);
        case 5: DD(5);
        case 6: DD(6);
        case 7: DD(7);
        case 8: DD(8);
        default: return WPC_EARG;
    }
#undef DD
}

int wpc_refine_prepare(int n, int nvec, const double* r, const double* norms, int mask, void* out,
                       int out_dtype, void* stream) {
    const long total = (long)n * nvec;
    if (total <= 0) return WPC_OK;
    if (nvec < 1 || nvec > 8) return WPC_EARG;
    const int grid = (int)((total + 255) / 256);
    hipStream_t s = as_stream(stream);
    if (out_dtype == WPC_F32)
        hipLaunchKernelGGL(refine_prepare_kernel<float>, dim3(grid), dim3(256), 0, s, total, nvec, r,
                           norms, mask, (float*)out);
    else if (out_dtype == WPC_F64)
        hipLaunchKernelGGL(refine_prepare_kernel<double>, dim3(grid), dim3(256), 0, s, total, nvec,
                           r, norms, mask, (double*)out);
    else
        return WPC_EDTYPE;
    return launch_status();
}

int wpc_refine_update(int n, int nvec, double* x, const void* c, int c_dtype, const double* norms,
                      int mask, double* c_out, void* stream) {
    const long total = (long)n * nvec;
    if (total <= 0) return WPC_OK;
    if (nvec < 1 || nvec > 8) return WPC_EARG;
    const int grid = (int)((total + 255) / 256);
    hipStream_t s = as_stream(stream);
    if (c_dtype == WPC_F32)
        hipLaunchKernelGGL(refine_update_kernel<float>, dim3(grid), dim3(256), 0, s, total, nvec, x,
                           (const float*)c, norms, mask, c_out);
    else if (c_dtype == WPC_F64)
        hipLaunchKernelGGL(refine_update_kernel<double>, dim3(grid), dim3(256), 0, s, total, nvec, x,
                           (const double*)c, norms, mask, c_out);
    else
        return WPC_EDTYPE;
    return launch_status();
}

int wpc_add_diagonal(int n, const int* offsets, const int* columns, double* values,
                     const double* d, double alpha, int* status, void* stream) {
    if (n <= 0) return WPC_OK;
    hipLaunchKernelGGL(add_diagonal_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
                       as_stream(stream), n, offsets, columns, values, d, alpha, status);
    return launch_status();
}

}  // extern "C"
