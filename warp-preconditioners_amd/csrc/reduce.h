// Deterministic block reductions shared by the streamed products and the
// Krylov kernels (256-thread workgroups).
#pragma once

#include "wpc_common.h"

namespace wpc {

constexpr int kStreamRows = 256;
constexpr int kXcds = 8;

// The dispatcher hands workgroups to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8), and each XCD has its own L2.  Renumbering the workgroups so that
// every XCD owns one contiguous range of logical blocks keeps the gathers of
// neighbouring rows (which share most of their columns) inside one L2.
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int x = b % kXcds, q = b / kXcds;
    const int per = nb / kXcds, extra = nb % kXcds;
    return x * per + min(x, extra) + q;
}

// 64-bit DPP move (both halves with the same lane pattern CTRL).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const long long hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((hi << 32) | lo);
}

// x of the lane whose index differs in bit log2(M) (M = 1 or 2, inside a quad):
// the DPP form of __shfl_xor(x, M) with no LDS round trip.
template <int M>
__device__ __forceinline__ double quad_xor(double x) {
    static_assert(M == 1 || M == 2, "quad strides only");
    return dpp_move<M == 1 ? 0xB1 : 0x4E>(x);
}

template <int M>
__device__ __forceinline__ float quad_xor(float x) {
    static_assert(M == 1 || M == 2, "quad strides only");
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(x), M == 1 ? 0xB1 : 0x4E, 0xF, 0xF, false));
}

__device__ __forceinline__ double read_lane(double x, int lane) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, lane);
    const long long hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double((hi << 32) | lo);
}

// Sum of x over the 64 lanes of a wavefront (all lanes active), the same value
// in every lane: pair and quad sums (DPP quad permutations), 8- and 16-lane rows
// (DPP half-row and row mirrors) -- VALU moves, no LDS round trips -- and the
// four row sums as (r0 + r1) + (r2 + r3).  Every step adds the same two values
// in both lanes it pairs, so all lanes agree bit for bit.
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp_move<0xB1>(x);   // quad_perm [1,0,3,2]
    x += dpp_move<0x4E>(x);   // quad_perm [2,3,0,1]
    x += dpp_move<0x141>(x);  // row_half_mirror
    x += dpp_move<0x140>(x);  // row_mirror
    return (read_lane(x, 0) + read_lane(x, 16)) + (read_lane(x, 32) + read_lane(x, 48));
}

constexpr int kBlockWaves = kStreamRows / kWave;

// Deterministic block sum of v[NV] (all NT threads) into v[0..NV) of thread 0:
// wave sums, then thread 0 adds the NT / 64 wave sums in wave order.
// red: kStreamRows doubles of scratch (NV * NT / 64 are used).
template <int NV, int NT = kStreamRows>
__device__ void block_sum(double* v, double* red) {
    constexpr int W = NT / kWave;
    static_assert(NV * W <= kStreamRows, "scratch too small");
    double s[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) s[q] = wave_sum(v[q]);
    __syncthreads();  // red may still be read by an earlier use
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) red[q * W + threadIdx.x / kWave] = s[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double t = 0.0;
            for (int w = 0; w < W; ++w) t += red[q * W + w];
            v[q] = t;
        }
    }
}

// Block partials are folded by a later launch (kernel boundaries order them): a
// device-scope fence inside a kernel writes back the whole L2 of the XCD on
// gfx950, which costs far more than one extra single-workgroup launch.
template <int NV>
__device__ __forceinline__ void write_partials(const double* v, double* partials,
                                               int block = -1) {
    if (threadIdx.x == 0) {
        const long b = block < 0 ? blockIdx.x : block;
#pragma unroll
        for (int q = 0; q < NV; ++q) partials[b * NV + q] = v[q];
    }
}

// The same two steps for a count nv <= 8 known only at run time (batched launches
// over systems with different right-hand-side counts); per vector they perform
// exactly the operations of block_sum<NV> / write_partials<NV>.
__device__ inline void block_sum_n(double* v, int nv, double* red) {
    double s[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] = q < nv ? wave_sum(v[q]) : 0.0;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q < nv) red[q * kBlockWaves + threadIdx.x / kWave] = s[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q >= nv) break;
            double t = 0.0;
            for (int w = 0; w < kBlockWaves; ++w) t += red[q * kBlockWaves + w];
            v[q] = t;
        }
    }
}

__device__ inline void write_partials_n(const double* v, int nv, double* partials, int block) {
    if (threadIdx.x == 0) {
        for (int q = 0; q < nv; ++q) partials[(long)block * nv + q] = v[q];
    }
}

// Fixed-order fold of partials[0..P) (stride nvec, vector v) by one workgroup;
// the result is returned to every thread.
__device__ inline double fold_block(const double* __restrict__ partials, int P, int nvec, int v,
                                    double* red) {
    double s = 0.0;
    for (int b = threadIdx.x; b < P; b += kStreamRows) s += partials[(long)b * nvec + v];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kStreamRows / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double out = red[0];
    __syncthreads();
    return out;
}

// Cross-wave end of the folds below: per-thread sums acc[0..nvec) -> out_s.
template <int NT>
__device__ void fold_reduce(const double* acc, int nvec, double* out_s) {
    __shared__ double red[NT / kWave][8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        if (v >= nvec) break;
        const double x = wave_sum(acc[v]);
        if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave][v] = x;
    }
    __syncthreads();
    if (threadIdx.x < nvec) {
        double t = 0.0;
        for (int w = 0; w < NT / kWave; ++w) t += red[w][threadIdx.x];
        out_s[threadIdx.x] = t;
    }
    __syncthreads();
}

// Fixed-order fold of partials[0..P) for all nvec (<= 8) vectors at once by a
// workgroup of NT threads; results land in out_s[0..nvec) (shared memory).
// Thread t sums rows t, t + NT, t + 2 NT, ... in that order.
template <int NT>
__device__ void fold_all(const double* __restrict__ partials, int P, int nvec, double* out_s) {
    double acc[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) acc[v] = 0.0;
    for (int b = threadIdx.x; b < P; b += NT)
        for (int v = 0; v < nvec; ++v) acc[v] += partials[(long)b * nvec + v];
    fold_reduce<NT>(acc, nvec, out_s);
}

// fold_all of two partial sets at once (results in out_a, out_b): each thread
// adds its rows of both in fold_all's order, with the loads of a row of each set
// issued together, and one cross-wavefront pass reduces all 2 nvec sums.
template <int NT>
__device__ void fold_two(const double* __restrict__ pa, int Pa, const double* __restrict__ pb,
                         int Pb, int nvec, double* out_a, double* out_b) {
    double acc[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0;
    const int P = Pa > Pb ? Pa : Pb;
    for (int b = threadIdx.x; b < P; b += NT) {
        double ta[8], tb[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            ta[v] = v < nvec && b < Pa ? pa[(long)b * nvec + v] : 0.0;
            tb[v] = v < nvec && b < Pb ? pb[(long)b * nvec + v] : 0.0;
        }
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            acc[v] += ta[v];  // (a row past P adds +0 to a sum that starts at +0)
            acc[8 + v] += tb[v];
        }
    }
    __shared__ double red[NT / kWave][16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        if ((v & 7) >= nvec) continue;
        const double x = wave_sum(acc[v]);
        if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave][v] = x;
    }
    __syncthreads();
    if (threadIdx.x < 2 * nvec) {
        const int v = threadIdx.x < nvec ? threadIdx.x : 8 + threadIdx.x - nvec;
        double t = 0.0;
        for (int w = 0; w < NT / kWave; ++w) t += red[w][v];
        (threadIdx.x < nvec ? out_a[threadIdx.x] : out_b[threadIdx.x - nvec]) = t;
    }
    __syncthreads();
}

// Partial counts P * NV up to this many are staged in registers by FoldRegs.
constexpr int kFoldStaged = 4096;

// The same fold split in two so that its loads overlap other work: load() issues
// every load of the thread's rows at once (one memory round trip instead of one
// per row), finish() adds them in fold_all's order (identical results) and
// reduces across the workgroup.  Rows beyond the staged capacity are read in
// finish(), so any P is folded correctly.
template <int NT, int NV>
struct FoldRegs {
    static constexpr int K = (kFoldStaged / NV + NT - 1) / NT;
    double val[K][NV];

    // Loads through a buffer descriptor whose range ends at row P: a row past P
    // reads 0 without a memory access.  The thread's rows within P go out at once,
    // branch-free per row (with the loads under a per-row condition the compiler
    // moved the first row's add into its branch and waited there, so the remaining
    // loads, and every load issued after them, went out one memory round trip
    // late); a whole group of rows past P (k NT >= P, uniform) is not issued.
    __device__ __forceinline__ void load(const double* __restrict__ partials, int P) {
        const __amdgpu_buffer_rsrc_t pr = __builtin_amdgcn_make_buffer_rsrc(
            (void*)partials, (short)0, P > 0 ? P * NV * (int)sizeof(double) : 0, kBufferWord3);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const unsigned row = threadIdx.x + k * NT;
            const bool issue = k == 0 || k * NT < P;
#pragma unroll
            for (int v = 0; v < NV; ++v)
                val[k][v] = issue ? buffer_ld<double>(pr, (row * NV + v) * (unsigned)sizeof(double))
                                  : 0.0;
        }
    }

    // Per-thread sums acc[0..NV) of the staged rows (and of any rows beyond the
    // staged capacity), in fold_all's order.  A row past P was loaded as +0, which
    // leaves the sum unchanged: it starts at +0 and so is never -0.
    __device__ __forceinline__ void sums(const double* __restrict__ partials, int P, double* acc) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += val[k][v];
        }
        for (int b = threadIdx.x + K * NT; b < P; b += NT)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[v] += partials[(long)b * NV + v];
    }

    __device__ __forceinline__ void finish(const double* __restrict__ partials, int P,
                                           double* out_s) {
        double acc[8];
        sums(partials, P, acc);
        fold_reduce<NT>(acc, NV, out_s);
    }
};

// State fields F0, F1 (, F2) of all nv vectors, held by the lanes of every
// wavefront (lane k * nv + v holds field k of vector v) and loaded with ordinary
// vector loads, so that they travel with a step's other loads (uniform reads
// would become scalar loads issued only after the rows have arrived).
template <int F0, int F1, int F2 = -1>
struct StateLanes {
    static constexpr int kFields = F2 < 0 ? 2 : 3;
    double val;

    __device__ __forceinline__ void load(const double* state, int nv) {
        const int lane = threadIdx.x & (kWave - 1);
        const int k = lane / nv, v = lane - k * nv;
        const int f = k == 0 ? F0 : (k == 1 ? F1 : F2);
        val = k < kFields ? state[f * nv + v] : 0.0;
    }

    __device__ __forceinline__ double get(int k, int nv, int v) const {
        const int lane = k * nv + v;
        const long long bits = __double_as_longlong(val);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)bits, lane);
        const long long hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
        return __longlong_as_double((hi << 32) | lo);
    }
};

}  // namespace wpc
