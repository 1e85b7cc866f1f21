// Construction of the sliced-ELL copy of a CSR matrix used by the sliced product
// mapping (kernels in stream_impl.h): slices of 64 rows, entries of a slice stored
// column-major, rows padded to the slice's longest row.  `sell_map` records, for
// every stored position, the CSR entry it copies (-1 for padding), so a matrix
// whose values change in place (a refit) refreshes its copy with one gather.
// An optional row order `rows` (a permutation of 0..n-1) puts row rows[i] in
// position i of the slices: ordering rows by length inside windows of a few
// hundred rows keeps the padding small when row lengths vary.

#include "wpc_common.h"

namespace wpc {

constexpr int kSlice = 64;
static_assert(kSlice == kWave, "a slice is one wavefront");

// slice_len[s] = longest row of slice s (one wavefront per slice).
__global__ __launch_bounds__(256) void sell_lengths_kernel(int n, int nslices,
                                                           const int* __restrict__ off,
                                                           const int* __restrict__ rows,
                                                           int* __restrict__ slice_len) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long row = idx < n ? (rows != nullptr ? rows[idx] : idx) : -1;
    int len = row >= 0 ? off[row + 1] - off[row] : 0;
    for (int m = kWave / 2; m > 0; m >>= 1) len = max(len, __shfl_xor(len, m, kWave));
    const long s = idx / kSlice;  // a wavefront is exactly one slice (kSlice == kWave)
    if ((threadIdx.x & (kWave - 1)) == 0 && s < nslices) slice_len[s] = len;
}

// Column-major slice entries; padding repeats the row's last column (or column 0
// for an empty row) with map -1.
__global__ __launch_bounds__(256) void sell_fill_kernel(int n, int nslices, const int* __restrict__ off,
                                                        const int* __restrict__ col,
                                                        const int* __restrict__ rows,
                                                        const int* __restrict__ slice_ptr,
                                                        const int* __restrict__ slice_len,
                                                        int* __restrict__ scol,
                                                        int* __restrict__ smap) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nslices * kSlice) return;
    const int s = (int)(idx / kSlice);
    const long base = (long)slice_ptr[s] * kSlice + (idx % kSlice);
    const int len = slice_len[s];
    const long row = idx < n ? (rows != nullptr ? rows[idx] : idx) : -1;
    const int a = row >= 0 ? off[row] : 0;
    const int b = row >= 0 ? off[row + 1] : 0;
    const int pad = b > a ? col[b - 1] : 0;
    for (int k = 0; k < len; ++k) {
        const long p = base + (long)k * kSlice;
        const bool real = a + k < b;
        scol[p] = real ? col[a + k] : pad;
        smap[p] = real ? a + k : -1;
    }
}

// int16 copy of the columns as differences col - row; overflow[0] counts entries
// whose difference does not fit (the int32 copy must then be used).  Padding
// repeats the row's last difference; an empty row pads with its own index (a
// valid column of a square matrix, row < ncol).
__global__ __launch_bounds__(256) void sell_col16_kernel(int n, int ncol, int nslices,
                                                         const int* __restrict__ off,
                                                         const int* __restrict__ rows,
                                                         const int* __restrict__ slice_ptr,
                                                         const int* __restrict__ slice_len,
                                                         const int* __restrict__ scol,
                                                         short* __restrict__ out, int* overflow) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nslices * kSlice) return;
    const int s = (int)(idx / kSlice);
    const long base = (long)slice_ptr[s] * kSlice + (idx % kSlice);
    const int len = slice_len[s];
    const long row = idx < n ? (rows != nullptr ? rows[idx] : idx) : -1;
    const int own = row >= 0 ? off[row + 1] - off[row] : 0;
    int bad = 0;
    long last = 0;
    if (own == 0 && row >= ncol) bad = 1;  // no column of its own to point at
    for (int k = 0; k < len; ++k) {
        const long p = base + (long)k * kSlice;
        long diff = last;
        if (k < own) {
            diff = (long)scol[p] - (row >= 0 ? row : 0);
            if (diff < -32768 || diff > 32767) {
                ++bad;
                diff = 0;
            }
            last = diff;
        }
        out[p] = (short)diff;
    }
    if (bad) atomicAdd(overflow, bad);
}

template <class E>
__global__ __launch_bounds__(256) void sell_values_kernel(long count, const int* __restrict__ smap,
                                                          const E* __restrict__ src,
                                                          E* __restrict__ dst) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long)gridDim.x * blockDim.x) {
        const int m = smap[i];
        dst[i] = m >= 0 ? src[m] : E(0);
    }
}

}  // namespace wpc

using namespace wpc;

extern "C" {

int wpc_sell_slices(int n) { return n > 0 ? (n + kSlice - 1) / kSlice : 0; }

int wpc_sell_lengths(int n, const int* offsets, const int* rows, int* slice_len, void* stream) {
    const int ns = wpc_sell_slices(n);
    if (ns == 0) return WPC_OK;
    hipLaunchKernelGGL(sell_lengths_kernel, dim3(blocks_for((long)ns * kSlice, 256)), dim3(256), 0,
                       as_stream(stream), n, ns, offsets, rows, slice_len);
    return launch_status();
}

int wpc_sell_fill(int n, const int* offsets, const int* columns, const int* rows,
                  const int* slice_ptr, const int* slice_len, int* sell_columns, int* sell_map,
                  void* stream) {
    const int ns = wpc_sell_slices(n);
    if (ns == 0) return WPC_OK;
    hipLaunchKernelGGL(sell_fill_kernel, dim3(blocks_for((long)ns * kSlice, 256)), dim3(256), 0,
                       as_stream(stream), n, ns, offsets, columns, rows, slice_ptr, slice_len,
                       sell_columns, sell_map);
    return launch_status();
}

int wpc_sell_col16(int n, int ncol, const int* offsets, const int* rows, const int* slice_ptr,
                   const int* slice_len, const int* sell_columns, short* sell_col16, int* overflow,
                   void* stream) {
    const int ns = wpc_sell_slices(n);
    if (ns == 0) return WPC_OK;
    hipLaunchKernelGGL(sell_col16_kernel, dim3(blocks_for((long)ns * kSlice, 256)), dim3(256), 0,
                       as_stream(stream), n, ncol, ns, offsets, rows, slice_ptr, slice_len,
                       sell_columns, sell_col16, overflow);
    return launch_status();
}

int wpc_sell_values(long count, const int* sell_map, const void* values, int elem_bytes,
                    void* sell_values, void* stream) {
    if (count <= 0) return WPC_OK;
    const int grid = (int)((count + 255) / 256 < 65536 ? (count + 255) / 256 : 65536);
    hipStream_t s = as_stream(stream);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(sell_values_kernel<uint64_t>, dim3(grid), dim3(256), 0, s, count,
                           sell_map, (const uint64_t*)values, (uint64_t*)sell_values);
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(sell_values_kernel<uint32_t>, dim3(grid), dim3(256), 0, s, count,
                           sell_map, (const uint32_t*)values, (uint32_t*)sell_values);
    else if (elem_bytes == 2)
        hipLaunchKernelGGL(sell_values_kernel<uint16_t>, dim3(grid), dim3(256), 0, s, count,
                           sell_map, (const uint16_t*)values, (uint16_t*)sell_values);
    else
        return WPC_EARG;
    return launch_status();
}

}  // extern "C"
