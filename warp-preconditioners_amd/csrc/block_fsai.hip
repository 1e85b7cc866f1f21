// Block-adaptive FSAI on b x b BSR matrices (benchmarks/block_fsai_candidate.py):
// one wavefront per block row.  Per growth step the equilibrated principal
// submatrix of the selected blocks is factored column by column (lanes over rows,
// each entry the same serial sum as the serial restatement), b identity
// right-hand sides are solved on b lanes, and the next block is the frontier block
// with the largest residual score (lanes over candidates, a wave argmax with ties
// to the smaller block).  The block row is then normalized by chol(Z_ii)^-1 and
// written as b padded scalar rows in pattern order.  All arithmetic follows the
// serial order, so the factor is bit-identical to it.

#include "wpc_common.h"

namespace wpc {

constexpr int kBfWaves = 4;       // block rows per workgroup
constexpr int kBfMaxWidth = 32;   // b * max_blocks
constexpr int kBfMaxBlocks = 16;

template <class T>
__device__ __forceinline__ T bsr_entry(const int* __restrict__ off, const int* __restrict__ col,
                                       const T* __restrict__ val, int b, int i, int j) {
    int lo = off[i / b], hi = off[i / b + 1];
    const int end = hi;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (col[mid] < j / b) lo = mid + 1;
        else hi = mid;
    }
    T v = T(0);
    if (lo < end && col[lo] == j / b) v = val[(long)lo * b * b + (i % b) * b + (j % b)];
    return v;
}

// block_fsai_candidate.py:45-59
template <class T>
__global__ __launch_bounds__(256) void block_diagonal_kernel(int n, int b, const int* __restrict__ off,
                                                             const int* __restrict__ col,
                                                             const T* __restrict__ val,
                                                             T* __restrict__ scale, int* status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T d = bsr_entry(off, col, val, b, i, i);
    if (d > T(0) && isfinite(d)) {
        scale[i] = T(1) / nsqrt(d);
    } else {
        atomicAdd(status, 1);
        scale[i] = T(0);
    }
}

// (score, block) ordering of the selection: larger score first, then smaller block;
// block < 0 means "none".
template <class T>
__device__ __forceinline__ bool better(T s, int c, T bs, int bc) {
    if (c < 0) return false;
    if (bc < 0) return true;
    return s > bs || (s == bs && c < bc);
}

// block_fsai_candidate.py:61-176
template <class T>
__global__ __launch_bounds__(kBfWaves * kWave) void block_fsai_kernel(
    int nb, int b, int B, const int* __restrict__ off, const int* __restrict__ col,
    const T* __restrict__ val, const T* __restrict__ scale, int* __restrict__ cols,
    T* __restrict__ vals, int* status, int refit) {
    __shared__ T chol_all[kBfWaves][kBfMaxWidth * kBfMaxWidth];
    __shared__ T z_all[kBfWaves][kBfMaxWidth * 4];
    __shared__ T norm_all[kBfWaves][16];
    __shared__ int pattern_all[kBfWaves][kBfMaxBlocks];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int i = blockIdx.x * kBfWaves + wave;
    if (i >= nb) return;  // the whole wavefront leaves together
    T* chol = chol_all[wave];
    T* z = z_all[wave];
    T* norm = norm_all[wave];
    int* pattern = pattern_all[wave];
    const int width = b * B;
    int size = 1;
    if (lane == 0) {
        pattern[0] = i;
        if (refit) {
            for (int p = 0; p < B; ++p) {
                const int c = cols[(long)i * b * width + p * b];
                if (c >= 0 && c / b != i) pattern[size++] = c / b;
            }
        }
    }
    size = __shfl(size, 0, kWave);
    wave_sync();
    for (int step = 0; step < B; ++step) {
        const int count = size * b;
        // Cholesky, column by column; lane t owns row p = q + t
        for (int q = 0; q < count; ++q) {
            const int p = q + lane;
            T v = T(0);
            if (p < count) {
                const int r = pattern[p / b] * b + p % b;
                const int c = pattern[q / b] * b + q % b;
                v = bsr_entry(off, col, val, b, r, c) * scale[r] * scale[c];
                for (int k = 0; k < q; ++k) v -= chol[p * kBfMaxWidth + k] * chol[q * kBfMaxWidth + k];
            }
            if (lane == 0) {
                if (v <= T(0) || !isfinite(v)) {
                    atomicAdd(status, 1);
                    v = T(1);
                }
                chol[q * kBfMaxWidth + q] = nsqrt(v);
            }
            wave_sync();
            if (lane > 0 && p < count) chol[p * kBfMaxWidth + q] = v / chol[q * kBfMaxWidth + q];
            wave_sync();
        }
        // b identity right-hand sides, one per lane
        if (lane < b) {
            const int c = lane;
            for (int p = 0; p < count; ++p) {
                T v = p == c ? T(1) : T(0);
                for (int q = 0; q < p; ++q) v -= chol[p * kBfMaxWidth + q] * z[q * 4 + c];
                z[p * 4 + c] = v / chol[p * kBfMaxWidth + p];
            }
            for (int rev = 0; rev < count; ++rev) {
                const int p = count - rev - 1;
                T v = z[p * 4 + c];
                for (int q = p + 1; q < count; ++q) v -= chol[q * kBfMaxWidth + p] * z[q * 4 + c];
                z[p * 4 + c] = v / chol[p * kBfMaxWidth + p];
            }
        }
        wave_sync();
        if (refit || step == B - 1) break;
        // score the strictly-lower frontier blocks of the selected block rows
        int total = 0;
        for (int p = 0; p < size; ++p) total += off[pattern[p] + 1] - off[pattern[p]];
        T best_score = T(0);
        int best = -1;
        for (int t = lane; t < total; t += kWave) {
            int p = 0, e = t;
            while (e >= off[pattern[p] + 1] - off[pattern[p]]) {
                e -= off[pattern[p] + 1] - off[pattern[p]];
                ++p;
            }
            const int candidate = col[off[pattern[p]] + e];
            bool selected = candidate >= i;
            for (int q = 0; q < size; ++q)
                if (pattern[q] == candidate) selected = true;
            if (selected) continue;
            T score = T(0);
            for (int r = 0; r < b; ++r) {
                const int cr = candidate * b + r;
                for (int c = 0; c < b; ++c) {
                    T residual = T(0);
                    for (int q = 0; q < count; ++q) {
                        const int j = pattern[q / b] * b + q % b;
                        residual += bsr_entry(off, col, val, b, cr, j) * scale[cr] * scale[j] *
                                    z[q * 4 + c];
                    }
                    score += residual * residual;
                }
            }
            if (score > T(0) && better(score, candidate, best_score, best)) {
                best_score = score;
                best = candidate;
            }
        }
        for (int m = kWave / 2; m > 0; m >>= 1) {
            const T os = shfl_xor(best_score, m, kWave);
            const int oc = __shfl_xor(best, m, kWave);
            if (better(os, oc, best_score, best)) {
                best_score = os;
                best = oc;
            }
        }
        if (best < 0) break;
        if (lane == 0) pattern[size] = best;
        ++size;
        wave_sync();
    }
    // normalize: Z <- Z chol(Z_ii)^-T
    if (lane == 0) {
        for (int r = 0; r < b; ++r) {
            for (int c = 0; c <= r; ++c) {
                T v = z[r * 4 + c];
                for (int k = 0; k < c; ++k) v -= norm[r * 4 + k] * norm[c * 4 + k];
                if (r == c) norm[r * 4 + c] = nsqrt(v);
                else norm[r * 4 + c] = v / norm[c * 4 + c];
            }
        }
    }
    wave_sync();
    for (int q = lane; q < size * b; q += kWave) {
        for (int r = 0; r < b; ++r) {
            T v = z[q * 4 + r];
            for (int k = 0; k < r; ++k) v -= norm[r * 4 + k] * z[q * 4 + k];
            z[q * 4 + r] = v / norm[r * 4 + r];
        }
    }
    wave_sync();
    for (int t = lane; t < b * width; t += kWave) {
        const int r = t / width, q = t % width;
        const long e = ((long)i * b + r) * width + q;
        int cj = -1;
        T vj = T(0);
        if (q < size * b) {
            cj = pattern[q / b] * b + q % b;
            vj = z[q * 4 + r] * scale[cj];
        }
        cols[e] = cj;
        vals[e] = vj;
    }
}

template <class T>
static int block_fsai_launch(int nb, int b, int B, const int* off, const int* col, const void* val,
                             void* scale, int* cols, void* vals, int* status, int refit,
                             hipStream_t s) {
    const int n = nb * b;
    hipLaunchKernelGGL(block_diagonal_kernel<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, b,
                       off, col, (const T*)val, (T*)scale, status);
    hipLaunchKernelGGL(block_fsai_kernel<T>, dim3(blocks_for(nb, kBfWaves)), dim3(kBfWaves * kWave),
                       0, s, nb, b, B, off, col, (const T*)val, (const T*)scale, cols, (T*)vals,
                       status, refit);
    return launch_status();
}

}  // namespace wpc

using namespace wpc;

extern "C" {

int wpc_block_fsai_build(int nb, int b, int max_blocks, const int* offsets, const int* columns,
                         const void* values, int dtype, void* scale, int* cols, void* vals,
                         int* status, int refit, void* stream) {
    if (nb <= 0) return WPC_OK;
    if (b < 2 || b > 4 || max_blocks < 1 || max_blocks > kBfMaxBlocks ||
        b * max_blocks > kBfMaxWidth)
        return WPC_EARG;
    hipStream_t s = as_stream(stream);
    if (dtype == WPC_F64)
        return block_fsai_launch<double>(nb, b, max_blocks, offsets, columns, values, scale, cols,
                                         vals, status, refit, s);
    if (dtype == WPC_F32)
        return block_fsai_launch<float>(nb, b, max_blocks, offsets, columns, values, scale, cols,
                                        vals, status, refit, s);
    return WPC_EDTYPE;
}

}  // extern "C"
