// Fixed-support FSAI refit on gfx950: topology checks, value maps and the per-row
// dense Cholesky refit (one wavefront per row, factor in LDS).
//
// The Cholesky row p is computed as a column sweep across lanes (lane q owns entry
// (p, q)), which applies the subtractions of each entry in the same ascending order
// as a serial left-looking factorisation; the forward solve is swept the same way;
// the back-solve is serial on lane 0.

#include "fsai_device.h"

namespace wpc {

__global__ __launch_bounds__(256) void compare_kernel(long count, const int* a, const int* b,
                                                      int* bad) {
    int local = 0;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (long)gridDim.x * blockDim.x)
        local += a[k] != b[k];
    if (local) atomicAdd(bad, local);
}

__device__ __forceinline__ int find(const int* off, const int* col, int i, int j) {
    int lo = off[i], hi = off[i + 1];
    const int end = hi;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return (lo < end && col[lo] == j) ? lo : -1;
}

// fsai_reuse.py:35-48 (source_map).
__global__ __launch_bounds__(256) void source_map_kernel(int n, const int* off, const int* col,
                                                         const int* soff, const int* scol, int b,
                                                         int* mapping) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = off[i]; e < off[i + 1]; ++e) {
        const int j = col[e];
        const int slot = find(soff, scol, i / b, j / b);
        mapping[e] = slot * b * b + (i % b) * b + j % b;
    }
}

// fsai_reuse.py:51-61 (transpose_map).
__global__ __launch_bounds__(256) void transpose_map_kernel(int n, const int* off, const int* col,
                                                            const int* toff, const int* tcol,
                                                            int* mapping) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = off[i]; e < off[i + 1]; ++e) mapping[e] = find(toff, tcol, col[e], i);
}

template <class E>
__global__ __launch_bounds__(256) void gather_kernel(long count, const E* src, const int* map,
                                                     E* out) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (long)gridDim.x * blockDim.x)
        out[e] = src[map[e]];
}

template <class E>
__global__ __launch_bounds__(256) void scatter_kernel(long count, const E* src, const int* map,
                                                      E* out) {
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (long)gridDim.x * blockDim.x)
        out[map[e]] = src[e];
}

constexpr int kRefitWaves = 4;

template <class T>
__host__ __device__ inline size_t refit_lds_bytes(int W) {
    size_t t = (size_t)(W + W + W * (W + 1) / 2) * sizeof(T);  // z, scl, chol
    size_t i = (size_t)W * sizeof(int);
    return (t + i + 15) & ~(size_t)15;
}

// fsai_reuse.py:76-134 (refit), one wavefront per row.
template <class S, class O>
__global__ __launch_bounds__(256) void refit_kernel(int n, const int* __restrict__ off,
                                                    const int* __restrict__ col,
                                                    const S* __restrict__ val,
                                                    const int* __restrict__ goff,
                                                    const int* __restrict__ gcol,
                                                    const S* __restrict__ scale,
                                                    typename Arith<S>::type floor, int W, O* out,
                                                    int* bad) {
    using T = typename Arith<S>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int i = blockIdx.x * kRefitWaves + wave;
    if (i >= n) return;
    unsigned char* base = smem + wave * refit_lds_bytes<T>(W);
    T* z = reinterpret_cast<T*>(base);
    T* scl = z + W;
    T* chol = scl + W;
    int* sup = reinterpret_cast<int*>(chol + W * (W + 1) / 2);

    const int start = goff[i];
    const int size = goff[i + 1] - start;
    if (lane < size) {
        sup[lane] = gcol[start + lane];
        scl[lane] = ldv<T, S>(scale, gcol[start + lane]);
    }
    wave_sync();
    bool valid = true;
    for (int p = 0; p < size; ++p) {
        const int r = sup[p];
        T a = T(0);
        if (lane <= p) {
            const int slot = find(off, col, r, sup[lane]);
            if (slot >= 0) a = ldv<T, S>(val, slot) * scl[p] * scl[lane];
        }
        for (int k = 0; k < p; ++k) {
            // entry (p, k) is final once columns < k have been subtracted
            const T lpk = shfl(a / chol[tri(k, k)], k);
            if (lane == k) a = lpk;
            // row p itself is still in registers: its (p, k) entry is lpk
            if (lane > k && lane <= p) a -= lpk * (lane == p ? lpk : chol[tri(lane, k)]);
        }
        if (lane == p) {
            if (!(a > floor) || !nfinite(a)) {
                valid = false;
                a = T(1);
            }
            a = nsqrt(a);
        }
        if (lane <= p) chol[tri(p, lane)] = a;
        wave_sync();
    }
    // forward solve L y = e_last (column sweep keeps each row's order)
    T y = T(0);
    if (lane < size && sup[lane] == i) y = T(1);
    for (int q = 0; q < size; ++q) {
        const T yq = shfl(y / chol[tri(q, q)], q);
        if (lane == q) y = yq;
        if (lane > q && lane < size) y -= chol[tri(lane, q)] * yq;
    }
    if (lane < size) z[lane] = y;
    wave_sync();
    if (lane == 0) {
        for (int rev = 0; rev < size; ++rev) {
            const int p = size - 1 - rev;
            T acc = z[p];
            for (int q = p + 1; q < size; ++q) acc -= chol[tri(q, p)] * z[q];
            z[p] = acc / chol[tri(p, p)];
        }
    }
    wave_sync();
    if (size > 0) {
        const T norm = nsqrt(z[size - 1]);
        if (lane < size) {
            const T value_t = z[lane] / norm * scl[lane];
            St<O>::template st<double>(out, start + lane, (double)value_t);
            const float stored = St<O>::template ld<float>(out, start + lane);
            const double stored_d = St<O>::template ld<double>(out, start + lane);
            if (!isfinite(stored_d)) valid = false;
            if (lane == size - 1 && !(stored > 0.0f)) valid = false;
        }
    }
    const bool any_bad = __ballot(!valid) != 0;
    if (any_bad && lane == 0) atomicAdd(bad, 1);
}

// Deterministic +-1 probe (splitmix64 of seed and index).
template <class T>
__global__ __launch_bounds__(256) void rademacher_kernel(int n, unsigned long long seed, T* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long x = seed * 0x9E3779B97F4A7C15ull + (unsigned long long)i + 1;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    out[i] = (x >> 63) ? T(1) : T(-1);
}

template <class S, class O>
static int refit_launch(int n, const int* off, const int* col, const void* val, const int* goff,
                        const int* gcol, const void* scale, typename Arith<S>::type floor, int W,
                        void* out, int* bad, hipStream_t s) {
    using T = typename Arith<S>::type;
    const size_t lds = kRefitWaves * refit_lds_bytes<T>(W);
    auto kern = refit_kernel<S, O>;
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return launch_status();
    }
    hipLaunchKernelGGL(kern, dim3(blocks_for(n, kRefitWaves)), dim3(kRefitWaves * kWave), lds, s, n,
                       off, col, (const S*)val, goff, gcol, (const S*)scale, floor, W, (O*)out, bad);
    return launch_status();
}

}  // namespace wpc

using namespace wpc;

namespace {
float round_half_value(double x) {
    if (x != x) return (float)x;
    double ax = x < 0 ? -x : x;
    if (ax == 0.0) return (float)x;
    if (ax >= 65520.0) return x < 0 ? -__builtin_inff() : __builtin_inff();
    int e;
    frexp(ax, &e);
    int ex = e - 1;
    if (ex < -14) ex = -14;
    double quantum = ldexp(1.0, ex - 10);
    double r = nearbyint(ax / quantum) * quantum;
    return (float)(x < 0 ? -r : r);
}
}  // namespace

extern "C" {

int wpc_compare_i32(long count, const int* a, const int* b, int* bad, void* stream) {
    if (count <= 0) return WPC_OK;
    long blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(compare_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                       count, a, b, bad);
    return launch_status();
}

int wpc_refit_source_map(int n, const int* offsets, const int* columns, const int* src_offsets,
                         const int* src_columns, int block, int* mapping, void* stream) {
    if (n <= 0) return WPC_OK;
    hipLaunchKernelGGL(source_map_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream),
                       n, offsets, columns, src_offsets, src_columns, block, mapping);
    return launch_status();
}

int wpc_refit_transpose_map(int n, const int* offsets, const int* columns, const int* t_offsets,
                            const int* t_columns, int* mapping, void* stream) {
    if (n <= 0) return WPC_OK;
    hipLaunchKernelGGL(transpose_map_kernel, dim3(blocks_for(n, 256)), dim3(256), 0,
                       as_stream(stream), n, offsets, columns, t_offsets, t_columns, mapping);
    return launch_status();
}

static unsigned stride_blocks(long count) {
    long b = (count + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b < 1 ? 1 : b);
}

int wpc_gather(long count, int elem_bytes, const void* src, const int* map, void* out,
               void* stream) {
    if (count <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    unsigned g = stride_blocks(count);
    switch (elem_bytes) {
        case 8: hipLaunchKernelGGL(gather_kernel<uint64_t>, dim3(g), dim3(256), 0, s, count, (const uint64_t*)src, map, (uint64_t*)out); break;
        case 4: hipLaunchKernelGGL(gather_kernel<uint32_t>, dim3(g), dim3(256), 0, s, count, (const uint32_t*)src, map, (uint32_t*)out); break;
        case 2: hipLaunchKernelGGL(gather_kernel<uint16_t>, dim3(g), dim3(256), 0, s, count, (const uint16_t*)src, map, (uint16_t*)out); break;
        default: return WPC_EARG;
    }
    return launch_status();
}

int wpc_scatter(long count, int elem_bytes, const void* src, const int* map, void* out,
                void* stream) {
    if (count <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    unsigned g = stride_blocks(count);
    switch (elem_bytes) {
        case 8: hipLaunchKernelGGL(scatter_kernel<uint64_t>, dim3(g), dim3(256), 0, s, count, (const uint64_t*)src, map, (uint64_t*)out); break;
        case 4: hipLaunchKernelGGL(scatter_kernel<uint32_t>, dim3(g), dim3(256), 0, s, count, (const uint32_t*)src, map, (uint32_t*)out); break;
        case 2: hipLaunchKernelGGL(scatter_kernel<uint16_t>, dim3(g), dim3(256), 0, s, count, (const uint16_t*)src, map, (uint16_t*)out); break;
        default: return WPC_EARG;
    }
    return launch_status();
}

int wpc_fsai_refit(int n, const int* offsets, const int* columns, const void* values, int dtype,
                   const int* g_offsets, const int* g_columns, const void* scale,
                   double pivot_floor, int width, void* out, int out_dtype, int* bad,
                   void* stream) {
    // LDS is sized for `width`, the largest stored row of G (max_row_size).
    const int W = width;
    if (W < 1 || W > 64) return WPC_EARG;
    if (n <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == WPC_F64 && out_dtype == WPC_F64)
        return refit_launch<double, double>(n, offsets, columns, values, g_offsets, g_columns, scale, pivot_floor, W, out, bad, s);
    if (dtype == WPC_F64 && out_dtype == WPC_F32)
        return refit_launch<double, float>(n, offsets, columns, values, g_offsets, g_columns, scale, pivot_floor, W, out, bad, s);
    if (dtype == WPC_F32 && out_dtype == WPC_F32)
        return refit_launch<float, float>(n, offsets, columns, values, g_offsets, g_columns, scale, (float)pivot_floor, W, out, bad, s);
    if (dtype == WPC_F16 && out_dtype == WPC_F16)
        return refit_launch<f16s, f16s>(n, offsets, columns, values, g_offsets, g_columns, scale, hnum::raw(round_half_value(pivot_floor)), W, out, bad, s);
    return WPC_EDTYPE;
}

int wpc_rademacher(int n, int seed, void* out, int dtype, void* stream) {
    if (n <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    unsigned g = blocks_for(n, 256);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(rademacher_kernel<double>, dim3(g), dim3(256), 0, s, n, (unsigned long long)seed, (double*)out);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(rademacher_kernel<float>, dim3(g), dim3(256), 0, s, n, (unsigned long long)seed, (float*)out);
    else
        return WPC_EDTYPE;
    return launch_status();
}

}  // extern "C"
