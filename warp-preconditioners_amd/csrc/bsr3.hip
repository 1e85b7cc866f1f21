// 3x3 block-sparse products on a sliced copy (configurations with 3 unknowns per
// vertex: elasticity, K (x) C): z = alpha * A x + beta * y with A in 3x3 BSR,
// vectors interleaved [scalar row][nvec] (scalar row 3i + c of block row i).
//
// Layout: block rows in slices of 64 (one wavefront), the block columns of a
// slice stored column-major exactly as the scalar sliced copy does (sell.hip
// builds them from the block offsets/columns: int32 with padding -1, or 16-bit
// offsets from a per-slice base with padding 0xFFFF), and each block's nine
// values stored component-major inside the slice: value q (row-major r * 3 + c)
// of the block in slot p = (slice_ptr[s] + k) * 64 + lane sits at
// ((slice_ptr[s] + k) * 9 + q) * 64 + lane, so every one of the nine loads of a
// group is one fully coalesced 64-lane stream.
//
// One thread per block row sums its blocks in storage order; per block it forms
// t_r = (v_r0 x_0 + v_r1 x_1) + v_r2 x_2 and adds it to the row's sum (the block
// product, then the accumulation, as Warp's bsr_mv adds A_ij x_j), so the sums
// are deterministic and equal to the oracle's bit for bit.  Gathers of the
// three x components of a block column go through a buffer descriptor (padding
// offsets fall outside it and read 0).  Optional fused w . z partials (one per
// workgroup and vector) as in the scalar products.

#include "reduce.h"
#include "wpc_common.h"

namespace wpc {
namespace {

// Blocks per round trip: as many as keep a group's values and gathered x within
// about 80 registers (9 values + 3 * NV gathered components per block).
template <class TV, class TX, int NV>
constexpr int kB3Group() {
    constexpr int regs = (9 * (int)sizeof(TV) + 3 * NV * (int)sizeof(TX)) / 4;
    constexpr int g = 80 / regs;
    return g < 1 ? 1 : (g > 4 ? 4 : g);
}

template <class TI> __device__ __forceinline__ unsigned b3_off(const TI* col, long p, unsigned stride) {
    return (unsigned)col[p] * stride;
}

// One group of G blocks at slice position k (m < G: a masked group whose entries
// past m re-read entry m - 1 and whose products are discarded by a select).
template <class TI, class TV, class TX, int NV, int G, bool FULL>
__device__ __forceinline__ void b3_group(TX (&acc)[3][NV], long sbase, int k, int lane,
                                         __amdgpu_buffer_rsrc_t xr, const TI* __restrict__ col,
                                         const TV* __restrict__ val, int m) {
    constexpr unsigned stride = 3 * NV * sizeof(TX);
    unsigned o[G];
    TX a[G][9];
#pragma unroll
    for (int j = 0; j < G; ++j) {
        const int e = FULL || j < m ? j : m - 1;
        const long slot = sbase + k + e;  // slice_ptr[s] + k + e
        o[j] = b3_off<TI>(col, slot * 64 + lane, stride);
#pragma unroll
        for (int q = 0; q < 9; ++q) a[j][q] = (TX)val[(slot * 9 + q) * 64 + lane];
    }
    TX g[G][3][NV];
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int v = 0; v < NV; ++v)
                g[j][c][v] = buffer_ld<TX>(xr, o[j] + (unsigned)((c * NV + v) * sizeof(TX)));
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                TX t = a[j][3 * r] * g[j][0][v];
                t += a[j][3 * r + 1] * g[j][1][v];
                t += a[j][3 * r + 2] * g[j][2][v];
                const TX sum = acc[r][v] + t;
                acc[r][v] = FULL || j < m ? sum : acc[r][v];
            }
        }
    }
}

template <class TI, class TV, class TX, int NV>
__device__ __forceinline__ void b3_accum(TX (&acc)[3][NV], long sbase, int len, int lane,
                                         __amdgpu_buffer_rsrc_t xr, const TI* __restrict__ col,
                                         const TV* __restrict__ val) {
    constexpr int G = kB3Group<TV, TX, NV>();
    int k = 0;
    for (; k + G <= len; k += G) b3_group<TI, TV, TX, NV, G, true>(acc, sbase, k, lane, xr, col, val, G);
    const int r = len - k;
    if constexpr (G > 1) {
        if (r > 0) b3_group<TI, TV, TX, NV, G, false>(acc, sbase, k, lane, xr, col, val, r);
    }
}

template <class TV, class TX, int NV, bool DOT>
__global__ __launch_bounds__(kStreamRows) void bsr3_sell_kernel(
    int nb, const int* __restrict__ slice_ptr, const int* __restrict__ slice_len,
    const int* __restrict__ slice_base, const void* __restrict__ col, int col16,
    const TV* __restrict__ val, const TX* __restrict__ x, const TX* y, TX* z, TX alpha, TX beta,
    const TX* __restrict__ w, double* __restrict__ partials) {
    __shared__ double red[kStreamRows];
    const int blk = xcd_block();
    const int idx = blk * kStreamRows + threadIdx.x;
    const bool live = idx < nb;
    const int lane = threadIdx.x & 63;
    // the wavefront's slice (uniform; scalar header loads), as in the scalar
    // sliced products: a wavefront past the last block row works on the last
    // slice and stores nothing
    const int wave0 = __builtin_amdgcn_readfirstlane(idx & ~63);
    const int slice = min(wave0, nb - 1) >> 6;
    const int sp = slice_ptr[slice];
    const int sl = slice_len[slice];
    const int sb = slice_base[slice];
    const int row = live ? idx : nb - 1;
    const bool use_a = alpha != TX(0);
    const int len = use_a && wave0 < nb ? sl : 0;
    constexpr unsigned stride = 3 * NV * sizeof(TX);
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(x + (long)sb * 3 * NV), (short)0, (int)(col16 ? 0xFFFFu * stride : kRange32),
        kBufferWord3);
    TX wv[3][NV];
    if (DOT) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int v = 0; v < NV; ++v) wv[r][v] = w[((long)row * 3 + r) * NV + v];
    }
    TX acc[3][NV];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[r][v] = TX(0);
    if (col16)
        b3_accum<unsigned short, TV, TX, NV>(acc, sp, len, lane, xr, (const unsigned short*)col, val);
    else
        b3_accum<int, TV, TX, NV>(acc, sp, len, lane, xr, (const int*)col, val);
    // every lane computes (a lane past nb on block row nb - 1) and only the
    // stores are predicated, as in the scalar sliced products
    double dot[NV];
    TX out[3][NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) dot[v] = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            TX t = TX(0);
            if (use_a) t = alpha * acc[r][v];
            if (beta != TX(0)) t += beta * y[((long)row * 3 + r) * NV + v];
            out[r][v] = t;
            if (DOT && live) dot[v] += (double)wv[r][v] * (double)t;
        }
    }
    if (DOT) {
#pragma unroll
        for (int v = 0; v < NV; ++v) asm volatile("" : "+v"(dot[v]));
    }
    if (live) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int v = 0; v < NV; ++v) z[((long)row * 3 + r) * NV + v] = out[r][v];
    }
    if (DOT) {
        block_sum<NV>(dot, red);
        write_partials<NV>(dot, partials, blk);
    }
}

// The nine values of each slot's block, component-major inside the slice (map:
// the block of each slot, -1 for padding, from wpc_sell_fill on the block pattern).
template <class E>
__global__ __launch_bounds__(256) void bsr3_values_kernel(long slots, const int* __restrict__ map,
                                                          const E* __restrict__ src,
                                                          E* __restrict__ dst) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < slots * 9;
         i += (long)gridDim.x * blockDim.x) {
        const long slot = i / 9;
        const int q = (int)(i % 9);
        const int m = map[slot];
        dst[((slot >> 6) * 9 + q) * 64 + (slot & 63)] = m >= 0 ? src[(long)m * 9 + q] : E(0);
    }
}

// out[v] = fixed-order fold of the P partials of vector v (one workgroup).
__global__ __launch_bounds__(kStreamRows) void b3_fold_kernel(int P, int nvec,
                                                              const double* __restrict__ partials,
                                                              double* __restrict__ out) {
    __shared__ double red[kStreamRows];
    for (int v = 0; v < nvec; ++v) {
        const double t = fold_block(partials, P, nvec, v, red);
        if (threadIdx.x == 0) out[v] = t;
    }
}

template <class TV, class TX, int NV>
int b3_launch(int nb, const int* sp, const int* sl, const int* sb, const void* col, int col16,
              const void* val, const void* x, const void* y, void* z, double alpha, double beta,
              const void* w, double* partials, hipStream_t st) {
    const int grid = (int)blocks_for(nb, kStreamRows);
    if (w != nullptr)
        hipLaunchKernelGGL((bsr3_sell_kernel<TV, TX, NV, true>), dim3(grid), dim3(kStreamRows), 0,
                           st, nb, sp, sl, sb, col, col16, (const TV*)val, (const TX*)x,
                           (const TX*)y, (TX*)z, (TX)alpha, (TX)beta, (const TX*)w, partials);
    else
        hipLaunchKernelGGL((bsr3_sell_kernel<TV, TX, NV, false>), dim3(grid), dim3(kStreamRows), 0,
                           st, nb, sp, sl, sb, col, col16, (const TV*)val, (const TX*)x,
                           (const TX*)y, (TX*)z, (TX)alpha, (TX)beta, (const TX*)w, partials);
    return launch_status();
}

int b3_fold(int grid, int nvec, const double* partials, double* out, hipStream_t st) {
    hipLaunchKernelGGL(b3_fold_kernel, dim3(1), dim3(kStreamRows), 0, st, grid, nvec, partials, out);
    return launch_status();
}

template <class TV, class TX>
int b3_nv(int nvec, int nb, const int* sp, const int* sl, const int* sb, const void* col,
          int col16, const void* val, const void* x, const void* y, void* z, double alpha,
          double beta, const void* w, double* partials, hipStream_t st) {
    switch (nvec) {
#define B3(N) case N: return b3_launch<TV, TX, N>(nb, sp, sl, sb, col, col16, val, x, y, z, alpha, beta, w, partials, st)
        B3(1); B3(2); B3(3); B3(4);
#undef B3
    }
    return WPC_EARG;
}

}  // namespace
}  // namespace wpc

using namespace wpc;

extern "C" {

int wpc_bsr3_sell_values(long slots, const int* sell_map, const void* values, int elem_bytes,
                         void* sell_values, void* stream) {
    if (slots <= 0) return WPC_OK;
    const long work = slots * 9;
    const int grid = (int)((work + 255) / 256 < 65536 ? (work + 255) / 256 : 65536);
    hipStream_t s = as_stream(stream);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(bsr3_values_kernel<uint64_t>, dim3(grid), dim3(256), 0, s, slots,
                           sell_map, (const uint64_t*)values, (uint64_t*)sell_values);
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(bsr3_values_kernel<uint32_t>, dim3(grid), dim3(256), 0, s, slots,
                           sell_map, (const uint32_t*)values, (uint32_t*)sell_values);
    else
        return WPC_EARG;
    return launch_status();
}

int wpc_bsr3_sell_stream(int nb, const int* slice_ptr, const int* slice_len,
                         const int* slice_base, const void* sell_columns, int col16,
                         const void* sell_values, int val_dtype, const void* x, const void* y,
                         void* z, int vec_dtype, int nvec, double alpha, double beta,
                         const void* w, double* partials, double* out, void* stream) {
    if (nb <= 0) return WPC_OK;
    if (!slice_ptr || !slice_len || !slice_base || !sell_columns || !sell_values) return WPC_EARG;
    if (nvec < 1 || nvec > 4) return WPC_EARG;
    if (w != nullptr && partials == nullptr) return WPC_EARG;
    hipStream_t s = as_stream(stream);
    int rc = WPC_EDTYPE;
    if (vec_dtype == WPC_F64 && val_dtype == WPC_F64)
        rc = b3_nv<double, double>(nvec, nb, slice_ptr, slice_len, slice_base, sell_columns, col16,
                                   sell_values, x, y, z, alpha, beta, w, partials, s);
    else if (vec_dtype == WPC_F64 && val_dtype == WPC_F32)
        rc = b3_nv<float, double>(nvec, nb, slice_ptr, slice_len, slice_base, sell_columns, col16,
                                  sell_values, x, y, z, alpha, beta, w, partials, s);
    else if (vec_dtype == WPC_F32 && val_dtype == WPC_F32)
        rc = b3_nv<float, float>(nvec, nb, slice_ptr, slice_len, slice_base, sell_columns, col16,
                                 sell_values, x, y, z, alpha, beta, w, partials, s);
    if (rc != WPC_OK || w == nullptr || out == nullptr) return rc;
    return b3_fold((int)blocks_for(nb, kStreamRows), nvec, partials, out, s);
}

}  // extern "C"
