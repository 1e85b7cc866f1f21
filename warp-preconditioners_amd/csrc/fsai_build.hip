// FSAI setup on gfx950: equilibration, the adaptive per-row build, packing,
// scalarisation, transposition and factor compression.
//
// The build runs one wavefront per matrix row.  Frontier candidates are scored
// lane-parallel; each score is the serial sum over the support in support order
// (so it is bit-identical to a serial evaluation), and the top-s selection is
// order independent (distinct columns by score desc, column asc), so lane-local
// top-s lists merged across the wave select exactly what a serial scan selects.
// The bordered Cholesky update runs as a column sweep across lanes, which keeps
// every row's subtraction order; the back-solve is serial on lane 0.  Per-wave
// state (support, packed Cholesky factor, z, v) lives in LDS.

#include "fsai_device.h"

namespace wpc {

// fsai.py:43-57 (diagonal).
template <class S>
__global__ __launch_bounds__(256) void diagonal_kernel(int n, const int* __restrict__ off,
                                                       const int* __restrict__ col,
                                                       const S* __restrict__ val, S* scale,
                                                       int* status) {
    using T = typename Arith<S>::type;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T d = entry<T, S>(off, col, val, i, i);
    if (d > T(0) && nfinite(d)) {
        stv<S, T>(scale, i, T(1) / nsqrt(d));
    } else {
        stv<S, T>(scale, i, T(0));
        atomicAdd(status, 1);
    }
}

template <class T>
__host__ __device__ inline size_t wave_lds_bytes(int W, int S) {
    size_t t = (size_t)(3 * W + W * (W + 1) / 2) * sizeof(T);
    size_t i = (size_t)(W + W + 1 + S) * sizeof(int);
    return (t + i + 15) & ~(size_t)15;
}

constexpr int kBuildWaves = 4;

// Total order used by the selection: larger score first, then smaller column.
template <class T>
__device__ __forceinline__ bool ahead(T s1, int c1, T s2, int c2) {
    return s1 > s2 || (s1 == s2 && c1 < c2);
}

// fsai.py:59-176 (build), one wavefront per row i.
template <class S, int SMAX>
__global__ __launch_bounds__(256) void build_kernel(
    int n, const int* __restrict__ off, const int* __restrict__ col, const S* __restrict__ val,
    const S* __restrict__ scale, typename Arith<S>::type tolerance,
    typename Arith<S>::type pivot_floor, int W, int step, int* __restrict__ cols_out,
    S* __restrict__ vals_out, int* status) {
    using T = typename Arith<S>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int i = blockIdx.x * kBuildWaves + wave;
    if (i >= n) return;  // whole wavefront leaves together; no block barriers below

    unsigned char* base = smem + wave * wave_lds_bytes<T>(W, SMAX);
    T* z = reinterpret_cast<T*>(base);
    T* v = z + W;
    T* scl = v + W;  // scale[pattern[p]]
    T* chol = scl + W;
    int* pattern = reinterpret_cast<int*>(chol + W * (W + 1) / 2);
    int* prefix = pattern + W;  // frontier prefix sums, W + 1 entries
    int* cand = prefix + W + 1;  // selected columns of one search, SMAX entries

    const T scale_i = ldv<T, S>(scale, i);
    if (lane == 0) {
        pattern[0] = i;
        chol[0] = T(1);
        z[0] = T(1);
        v[0] = T(1);
        scl[0] = scale_i;
        prefix[0] = 0;
    }
    wave_sync();

    int size = 1;
    while (size < W) {
        // ---- frontier prefix sums over the support rows
        int len = 0;
        if (lane < size) {
            const int r = pattern[lane];
            len = off[r + 1] - off[r];
        }
        int pre = len;
        for (int d = 1; d < kWave; d <<= 1) {
            const int o = __shfl_up(pre, d, kWave);
            if (lane >= d) pre += o;
        }
        if (lane < size) prefix[lane + 1] = pre;
        const int total = shfl(pre, size - 1);
        wave_sync();

        // ---- lane-parallel scoring into lane-local top-s lists
        int lc[SMAX];
        T ls[SMAX];
#pragma unroll
        for (int r = 0; r < SMAX; ++r) {
            lc[r] = -1;
            ls[r] = T(0);
        }
        for (int t = lane; t < total; t += kWave) {
            int lo = 0, hi = size;  // largest p with prefix[p] <= t
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (prefix[mid] <= t) lo = mid;
                else hi = mid;
            }
            const int row = pattern[lo];
            const int c = col[off[row] + (t - prefix[lo])];
            bool skip = c >= i;
            for (int q = 0; q < size && !skip; ++q) skip = pattern[q] == c;
#pragma unroll
            for (int r = 0; r < SMAX; ++r) skip = skip || lc[r] == c;
            if (skip) continue;
            const T sc = ldv<T, S>(scale, c);
            T residual = T(0);
            for (int q0 = 0; q0 < size; q0 += 4) {
                // A[c, pattern[q0..q0+4)] by lockstep searches, summed in q order
                // (4 at a time: 8 needed registers that cost more than the round
                // trips they saved)
                T e[4];
                entries<T, S, 4>(off, col, val, c, pattern + q0, min(4, size - q0), e);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q0 + q < size) residual += e[q] * sc * scl[q0 + q] * z[q0 + q];
            }
            const T score = nabs(residual);
            int pos = -1;
#pragma unroll
            for (int r = 0; r < SMAX; ++r)
                if (pos < 0 && r < step && (score > ls[r] || (score == ls[r] && c < lc[r]))) pos = r;
            if (pos >= 0) {
#pragma unroll
                for (int r = SMAX - 1; r > 0; --r)
                    if (r > pos && r < step) {
                        lc[r] = lc[r - 1];
                        ls[r] = ls[r - 1];
                    }
#pragma unroll
                for (int r = 0; r < SMAX; ++r)
                    if (r == pos) {
                        lc[r] = c;
                        ls[r] = score;
                    }
            }
        }

        // ---- merge lane lists: `step` rounds of a wave-wide argmax
        int head = 0;
        int count = 0;
        for (int round = 0; round < step; ++round) {
            int hc = -1;
            T hs = T(0);
#pragma unroll
            for (int r = 0; r < SMAX; ++r)
                if (r == head) {
                    hc = lc[r];
                    hs = ls[r];
                }
            if (head >= step) hc = -1;
            int bc = hc;
            T bs = hs;
            for (int m = kWave >> 1; m > 0; m >>= 1) {
                const int oc = __shfl_xor(bc, m, kWave);
                const T os = shfl(bs, lane ^ m);
                if (oc >= 0 && (bc < 0 || ahead(os, oc, bs, bc))) {
                    bc = oc;
                    bs = os;
                }
            }
            if (bc < 0) break;
            if (lane == 0) cand[round] = bc;
            if (hc == bc) ++head;
            ++count;
        }
        if (count == 0) break;
        wave_sync();

        // ---- border the Cholesky factor for each selected candidate
        const T old_z0 = z[0];
        bool stopped = false;
        for (int rank = 0; rank < count; ++rank) {
            if (size == W) break;
            const int best = cand[rank];
            const T sb = ldv<T, S>(scale, best);
            T a = T(0);
            if (lane < size) a = entry<T, S>(off, col, val, best, pattern[lane]) * sb * scl[lane];
            T pivot = T(1);
            T wl = T(0);  // this lane's w[lane]
            for (int q = 0; q < size; ++q) {
                const T wq = shfl(a / chol[tri(q, q)], q);
                if (lane == q) wl = wq;
                if (lane > q && lane < size) a -= chol[tri(lane, q)] * wq;
                pivot -= wq * wq;
            }
            if (!(pivot > pivot_floor) || !nfinite(pivot)) {
                if (lane == 0) atomicAdd(status + 1, 1);
                stopped = true;
                break;
            }
            const T root = nsqrt(pivot);
            const T prod = lane < size ? wl * v[lane] : T(0);
            T rhs = T(0);
            for (int p = 0; p < size; ++p) rhs -= shfl(prod, p);
            if (lane < size) chol[tri(size, lane)] = wl;
            if (lane == 0) {
                chol[tri(size, size)] = root;
                pattern[size] = best;
                scl[size] = sb;
                v[size] = rhs / root;
            }
            ++size;
            wave_sync();
        }

        // ---- back-solve A[S,S] z = e_0 (serial, reference subtraction order)
        if (lane == 0) {
            for (int rev = 0; rev < size; ++rev) {
                const int p = size - 1 - rev;
                T acc = v[p];
                for (int q = p + 1; q < size; ++q) acc -= chol[tri(q, p)] * z[q];
                z[p] = acc / chol[tri(p, p)];
            }
        }
        wave_sync();
        if (stopped || (tolerance > T(0) && T(1) - old_z0 / z[0] <= tolerance)) break;
    }

    // ---- emit G[i, S] = z / sqrt(z0) * scale[S], padded with -1 / 0
    const T norm = nsqrt(z[0]);
    bool invalid = false;
    for (int p = lane; p < W; p += kWave) {
        const long e = (long)i * W + p;
        if (p < size) {
            const T value = z[p] / norm * scl[p];
            cols_out[e] = pattern[p];
            stv<S, T>(vals_out, e, value);
            if (!nfinite(value)) invalid = true;
            if (p == 0 && value <= T(0)) invalid = true;
        } else {
            cols_out[e] = -1;
            stv<S, T>(vals_out, e, T(0));
        }
    }
    if (__ballot(invalid) != 0 && lane == 0) atomicAdd(status + 2, 1);
}

template <class S, int SMAX>
static int build_launch(int n, const int* off, const int* col, const void* val, const void* scale,
                        typename Arith<S>::type tol, typename Arith<S>::type floor, int W, int step,
                        int* cols_out, void* vals_out, int* status, hipStream_t s) {
    using T = typename Arith<S>::type;
    const size_t lds = kBuildWaves * wave_lds_bytes<T>(W, SMAX);
    if (lds > 160 * 1024) return WPC_EARG;
    auto kern = build_kernel<S, SMAX>;
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return launch_status();
    }
    unsigned grid = blocks_for(n, kBuildWaves);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBuildWaves * kWave), lds, s, n, off, col,
                       (const S*)val, (const S*)scale, tol, floor, W, step, cols_out, (S*)vals_out,
                       status);
    return launch_status();
}

template <class S>
static int build_dispatch(int n, const int* off, const int* col, const void* val,
                          const void* scale, typename Arith<S>::type tol,
                          typename Arith<S>::type floor, int W, int step, int* cols_out,
                          void* vals_out, int* status, hipStream_t s) {
#define B(SM) return build_launch<S, SM>(n, off, col, val, scale, tol, floor, W, step, cols_out, vals_out, status, s)
    if (step <= 1) B(1);
    if (step <= 2) B(2);
    if (step <= 4) B(4);
    if (step <= 8) B(8);
    if (step <= 16) B(16);
    if (step <= 32) B(32);
    B(64);
#undef B
}

// ---------------------------------------------------------------- packing
// fsai_pack.py:15-39 (sort_rows).
template <class E>
__global__ __launch_bounds__(256) void sort_rows_kernel(int n, int W, int* columns, E* values,
                                                        int* counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) counts[0] = 0;
    if (i >= n) return;
    int* c = columns + (long)i * W;
    E* v = values + (long)i * W;
    int size = 0;
    for (int p = 0; p < W; ++p) {
        const int cp = c[p];
        if (cp < 0) continue;
        const E vp = v[p];
        int q = size;
        while (q > 0 && c[q - 1] > cp) {
            c[q] = c[q - 1];
            v[q] = v[q - 1];
            --q;
        }
        c[q] = cp;
        v[q] = vp;
        ++size;
    }
    for (int p = size; p < W; ++p) c[p] = -1;
    counts[i + 1] = size;
}

// One-workgroup inclusive scan (setup only; wp.utils.array_scan).  Tiles of 8192
// entries: coalesced loads through LDS, each thread scans 8 consecutive entries,
// the 1024 thread totals are scanned across the workgroup (wavefront shuffles, then
// the 16 wavefront totals), and the results leave through LDS again, coalesced.
constexpr int kScanThreads = 1024, kScanPer = 8, kScanTile = kScanThreads * kScanPer;

__global__ __launch_bounds__(kScanThreads) void scan_kernel(int count, const int* in, int* out) {
    __shared__ int tile[kScanTile];
    __shared__ int wave_tot[kScanThreads / kWave];
    __shared__ int carry_s;
    const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
    int carry = 0;
    for (int base = 0; base < count; base += kScanTile) {
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const int idx = base + k * kScanThreads + t;
            tile[k * kScanThreads + t] = idx < count ? in[idx] : 0;
        }
        __syncthreads();
        int v[kScanPer];
        int run = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            run += tile[t * kScanPer + k];
            v[k] = run;
        }
        int pre = run;  // inclusive scan of the thread totals within the wavefront
        for (int d = 1; d < kWave; d <<= 1) {
            const int o = __shfl_up(pre, d, kWave);
            if (lane >= d) pre += o;
        }
        if (lane == kWave - 1) wave_tot[wave] = pre;
        __syncthreads();
        int before = carry + pre - run;  // everything ahead of this thread's entries
        for (int w = 0; w < wave; ++w) before += wave_tot[w];
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) tile[t * kScanPer + k] = v[k] + before;
        if (t == kScanThreads - 1) carry_s = v[kScanPer - 1] + before;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const int idx = base + k * kScanThreads + t;
            if (idx < count) out[idx] = tile[k * kScanThreads + t];
        }
        carry = carry_s;
        __syncthreads();
    }
}

// fsai_pack.py:41-53 (pack).
template <class E>
__global__ __launch_bounds__(256) void pack_kernel(int n, int W, const int* offsets,
                                                   const int* columns, const E* values,
                                                   int* outcols, E* outvals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = offsets[i];
    const int cnt = offsets[i + 1] - b;
    for (int p = 0; p < cnt; ++p) {
        outcols[b + p] = columns[(long)i * W + p];
        outvals[b + p] = values[(long)i * W + p];
    }
}

// fsai_pack.py:93-125 (scalarize).
template <class E>
__global__ __launch_bounds__(256) void scalarize_kernel(int nb, int b, const int* offsets,
                                                        const int* columns, const E* values,
                                                        int* outoff, int* outcol, E* outval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = nb * b;
    if (i >= n) return;
    const int r = i / b, component = i % b;
    const int start = offsets[r];
    const int count = offsets[r + 1] - start;
    const long base = (long)start * b * b + (long)component * count * b;
    outoff[i] = (int)base;
    for (int p = 0; p < count; ++p)
        for (int c = 0; c < b; ++c) {
            const long dest = base + (long)p * b + c;
            outcol[dest] = columns[start + p] * b + c;
            outval[dest] = values[(long)(start + p) * b * b + component * b + c];
        }
    if (i == n - 1) outoff[n] = (int)(base + (long)count * b);
}

// ---------------------------------------------------------------- transpose
__global__ __launch_bounds__(256) void count_columns_kernel(int nnz, const int* col, int* counts) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += gridDim.x * blockDim.x)
        atomicAdd(counts + col[e] + 1, 1);
}

template <class E>
__global__ __launch_bounds__(256) void transpose_fill_kernel(int nrow, const int* off,
                                                             const int* col, const E* val,
                                                             int* cursor, int* tcol, E* tval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    for (int e = off[i]; e < off[i + 1]; ++e) {
        const int p = atomicAdd(cursor + col[e], 1);
        tcol[p] = i;
        tval[p] = val[e];
    }
}

// Atomic placement leaves each output row unordered; sort rows by column.
template <class E>
__global__ __launch_bounds__(256) void sort_segments_kernel(int nrow, const int* off, int* col,
                                                            E* val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    const int b = off[i], e = off[i + 1];
    for (int k = b + 1; k < e; ++k) {
        const int ck = col[k];
        const E vk = val[k];
        int q = k;
        while (q > b && col[q - 1] > ck) {
            col[q] = col[q - 1];
            val[q] = val[q - 1];
            --q;
        }
        col[q] = ck;
        val[q] = vk;
    }
}

__global__ __launch_bounds__(256) void zero_i32_kernel(int count, int* dst) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x)
        dst[k] = 0;
}

__global__ __launch_bounds__(256) void copy_i32_kernel(int count, const int* src, int* dst) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x)
        dst[k] = src[k];
}

template <class E>
static int transpose_launch(int nrow, int ncol, int nnz, const int* off, const int* col,
                            const void* val, int* toff, int* tcol, void* tval, int* work,
                            hipStream_t s) {
    // work: counts[ncol + 1] -> scanned into toff; then reused as the cursor.
    hipLaunchKernelGGL(zero_i32_kernel, dim3(blocks_for(ncol + 1, 256)), dim3(256), 0, s, ncol + 1,
                       work);
    if (nnz > 0)
        hipLaunchKernelGGL(count_columns_kernel, dim3(blocks_for(nnz, 256)), dim3(256), 0, s, nnz,
                           col, work);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, s, ncol + 1, work, toff);
    hipLaunchKernelGGL(copy_i32_kernel, dim3(blocks_for(ncol + 1, 256)), dim3(256), 0, s, ncol + 1,
                       toff, work);
    if (nrow > 0) {
        hipLaunchKernelGGL((transpose_fill_kernel<E>), dim3(blocks_for(nrow, 256)), dim3(256), 0,
                           s, nrow, off, col, (const E*)val, work, tcol, (E*)tval);
    }
    if (ncol > 0)
        hipLaunchKernelGGL((sort_segments_kernel<E>), dim3(blocks_for(ncol, 256)), dim3(256), 0, s,
                           ncol, toff, tcol, (E*)tval);
    return launch_status();
}

// ---------------------------------------------------------------- validation
// fsai.py:181-196 (_validate_float32_factor).
template <class S>
__global__ __launch_bounds__(256) void validate_kernel(int n, const int* off, const int* col,
                                                       const S* val, int* status) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    bool invalid = false;
    for (int e = off[row]; e < off[row + 1]; ++e) {
        const float v = St<S>::template ld<float>(val, e);
        if (!isfinite(v)) invalid = true;
        if (col[e] == row && !(v > 0.0f)) invalid = true;
    }
    if (invalid) atomicAdd(status, 1);
}

// Row-scaled narrow factor (no reference counterpart; see wpc.h).
template <class S, class Q>
__global__ __launch_bounds__(256) void compress_scaled_kernel(int n, const int* off,
                                                              const S* val, Q* q, float* scale2,
                                                              int* status) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int b = off[row], e = off[row + 1];
    double m = 0.0;
    bool invalid = false;
    for (int k = b; k < e; ++k) {
        const double a = St<S>::template ld<double>(val, k);
        if (!isfinite(a)) invalid = true;
        m = fmax(m, fabs(a));
    }
    if (!(m > 0.0) || !isfinite(m)) {
        invalid = true;
        m = 1.0;
    }
    const double inv = 1.0 / m;
    for (int k = b; k < e; ++k) {
        const double a = St<S>::template ld<double>(val, k) * inv;
        St<Q>::template st<double>(q, k, a);
        if (k == e - 1 && !(St<Q>::template ld<float>(q, k) > 0.0f)) invalid = true;
    }
    if (e == b) invalid = true;
    scale2[row] = (float)(m * m);
    if (invalid) atomicAdd(status, 1);
}

}  // namespace wpc

using namespace wpc;

namespace {

// Host-side IEEE binary16 rounding of a double (ties to even), returned as float.
float round_half_host(double x) {
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

int wpc_fsai_diagonal(int n, const int* offsets, const int* columns, const void* values, int dtype,
                      void* scale, int* status, void* stream) {
    if (n <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    unsigned g = blocks_for(n, 256);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(diagonal_kernel<double>, dim3(g), dim3(256), 0, s, n, offsets, columns,
                           (const double*)values, (double*)scale, status);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(diagonal_kernel<float>, dim3(g), dim3(256), 0, s, n, offsets, columns,
                           (const float*)values, (float*)scale, status);
    else if (dtype == WPC_F16)
        hipLaunchKernelGGL(diagonal_kernel<f16s>, dim3(g), dim3(256), 0, s, n, offsets, columns,
                           (const f16s*)values, (f16s*)scale, status);
    else
        return WPC_EDTYPE;
    return launch_status();
}

int wpc_fsai_build(int n, const int* offsets, const int* columns, const void* values, int dtype,
                   const void* scale, double tolerance, double pivot_floor, int width, int step,
                   int* cols_out, void* vals_out, int* status, void* stream) {
    if (width < 1 || width > 64 || step < 1 || step > 64) return WPC_EARG;
    if (n <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    if (dtype == WPC_F64)
        return build_dispatch<double>(n, offsets, columns, values, scale, tolerance, pivot_floor,
                                      width, step, cols_out, vals_out, status, s);
    if (dtype == WPC_F32)
        return build_dispatch<float>(n, offsets, columns, values, scale, (float)tolerance,
                                     (float)pivot_floor, width, step, cols_out, vals_out, status, s);
    if (dtype == WPC_F16) {
        // hnum::raw takes an already-rounded binary16 value (rounded here, on the host).
        return build_dispatch<f16s>(n, offsets, columns, values, scale,
                                    hnum::raw(round_half_host(tolerance)),
                                    hnum::raw(round_half_host(pivot_floor)), width, step, cols_out,
                                    vals_out, status, s);
    }
    return WPC_EDTYPE;
}

int wpc_fsai_sort_rows(int n, int width, int* columns, void* values, int dtype, int* counts,
                       void* stream) {
    hipStream_t s = as_stream(stream);
    unsigned g = blocks_for(n > 0 ? n : 1, 256);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(sort_rows_kernel<double>, dim3(g), dim3(256), 0, s, n, width, columns, (double*)values, counts);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(sort_rows_kernel<float>, dim3(g), dim3(256), 0, s, n, width, columns, (float*)values, counts);
    else if (dtype == WPC_F16 || dtype == WPC_BF16)
        hipLaunchKernelGGL(sort_rows_kernel<uint16_t>, dim3(g), dim3(256), 0, s, n, width, columns, (uint16_t*)values, counts);
    else
        return WPC_EDTYPE;
    return launch_status();
}

int wpc_scan_inclusive(int count, const int* in, int* out, void* stream) {
    if (count <= 0) return WPC_OK;
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, as_stream(stream), count, in, out);
    return launch_status();
}

int wpc_fsai_pack(int n, int width, const int* offsets, const int* columns, const void* values,
                  int dtype, int* out_columns, void* out_values, void* stream) {
    if (n <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    unsigned g = blocks_for(n, 256);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(pack_kernel<double>, dim3(g), dim3(256), 0, s, n, width, offsets, columns, (const double*)values, out_columns, (double*)out_values);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(pack_kernel<float>, dim3(g), dim3(256), 0, s, n, width, offsets, columns, (const float*)values, out_columns, (float*)out_values);
    else if (dtype == WPC_F16 || dtype == WPC_BF16)
        hipLaunchKernelGGL(pack_kernel<uint16_t>, dim3(g), dim3(256), 0, s, n, width, offsets, columns, (const uint16_t*)values, out_columns, (uint16_t*)out_values);
    else
        return WPC_EDTYPE;
    return launch_status();
}

int wpc_bsr_scalarize(int nb, int b, const int* offsets, const int* columns, const void* values,
                      int dtype, int* out_offsets, int* out_columns, void* out_values,
                      void* stream) {
    hipStream_t s = as_stream(stream);
    const int n = nb * b;
    if (n <= 0) {
        hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(256), 0, s, 1, out_offsets);
        return launch_status();
    }
    unsigned g = blocks_for(n, 256);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(scalarize_kernel<double>, dim3(g), dim3(256), 0, s, nb, b, offsets, columns, (const double*)values, out_offsets, out_columns, (double*)out_values);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(scalarize_kernel<float>, dim3(g), dim3(256), 0, s, nb, b, offsets, columns, (const float*)values, out_offsets, out_columns, (float*)out_values);
    else if (dtype == WPC_F16 || dtype == WPC_BF16)
        hipLaunchKernelGGL(scalarize_kernel<uint16_t>, dim3(g), dim3(256), 0, s, nb, b, offsets, columns, (const uint16_t*)values, out_offsets, out_columns, (uint16_t*)out_values);
    else
        return WPC_EDTYPE;
    return launch_status();
}

int wpc_csr_transpose(int nrow, int ncol, int nnz, const int* offsets, const int* columns,
                      const void* values, int elem_bytes, int* t_offsets, int* t_columns,
                      void* t_values, int* workspace, void* stream) {
    hipStream_t s = as_stream(stream);
    if (ncol < 0 || nrow < 0) return WPC_EARG;
    switch (elem_bytes) {
        case 8: return transpose_launch<uint64_t>(nrow, ncol, nnz, offsets, columns, values, t_offsets, t_columns, t_values, workspace, s);
        case 4: return transpose_launch<uint32_t>(nrow, ncol, nnz, offsets, columns, values, t_offsets, t_columns, t_values, workspace, s);
        case 2: return transpose_launch<uint16_t>(nrow, ncol, nnz, offsets, columns, values, t_offsets, t_columns, t_values, workspace, s);
    }
    return WPC_EARG;
}

int wpc_fsai_validate_factor(int n, const int* offsets, const int* columns, const void* values,
                             int dtype, int* status, void* stream) {
    if (n <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    unsigned g = blocks_for(n, 256);
    if (dtype == WPC_F64)
        hipLaunchKernelGGL(validate_kernel<double>, dim3(g), dim3(256), 0, s, n, offsets, columns, (const double*)values, status);
    else if (dtype == WPC_F32)
        hipLaunchKernelGGL(validate_kernel<float>, dim3(g), dim3(256), 0, s, n, offsets, columns, (const float*)values, status);
    else if (dtype == WPC_F16)
        hipLaunchKernelGGL(validate_kernel<f16s>, dim3(g), dim3(256), 0, s, n, offsets, columns, (const f16s*)values, status);
    else if (dtype == WPC_BF16)
        hipLaunchKernelGGL(validate_kernel<bf16s>, dim3(g), dim3(256), 0, s, n, offsets, columns, (const bf16s*)values, status);
    else
        return WPC_EDTYPE;
    return launch_status();
}

int wpc_fsai_compress_scaled(int n, const int* offsets, const int* columns, const void* values,
                             int dtype, void* q_values, int q_dtype, float* row_scale2,
                             int* status, void* stream) {
    (void)columns;
    if (n <= 0) return WPC_OK;
    hipStream_t s = as_stream(stream);
    unsigned g = blocks_for(n, 256);
#define CS(SS, QQ) do { hipLaunchKernelGGL((compress_scaled_kernel<SS, QQ>), dim3(g), dim3(256), 0, s, n, offsets, (const SS*)values, (QQ*)q_values, row_scale2, status); return launch_status(); } while (0)
    if (dtype == WPC_F64) {
        if (q_dtype == WPC_F16) CS(double, f16s);
        if (q_dtype == WPC_BF16) CS(double, bf16s);
        if (q_dtype == WPC_F32) CS(double, float);
    } else if (dtype == WPC_F32) {
        if (q_dtype == WPC_F16) CS(float, f16s);
        if (q_dtype == WPC_BF16) CS(float, bf16s);
    }
#undef CS
    return WPC_EDTYPE;
}

}  // extern "C"
