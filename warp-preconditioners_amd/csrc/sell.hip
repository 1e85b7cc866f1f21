// Construction of the sliced-ELL copy of a CSR matrix used by the sliced product
// mapping (kernels in stream_impl.h): slices of 64 rows, entries of a slice stored
// column-major, rows padded to the slice's longest row.  `sell_map` records, for
// every stored position, the CSR entry it copies (-1 for padding), so a matrix
// whose values change in place (a refit) refreshes its copy with one gather.
// An optional row order `rows` (a permutation of 0..n-1) puts row rows[i] in
// position i of the slices: ordering rows by length inside windows of a few
// hundred rows keeps the padding small when row lengths vary.

#include <climits>

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

// Column-major slice entries; a padding entry stores column -1 and map -1.  The
// products gather through a buffer descriptor whose range a padding entry's
// offset lies outside of: such a load returns 0 without a memory access, and the
// padding value is 0, so padding adds exactly +0 whatever x holds (inf, NaN).
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
    for (int k = 0; k < len; ++k) {
        const long p = base + (long)k * kSlice;
        const bool real = a + k < b;
        scol[p] = real ? col[a + k] : -1;
        smap[p] = real ? a + k : -1;
    }
}

// 16-bit copy of the columns: slice_base[s] = the smallest column of slice s, and
// each entry stores col - slice_base[s] as an unsigned 16-bit offset; 0xFFFF marks
// padding (the products' descriptor spans offsets 0..0xFFFE).  overflow[0] counts
// slices whose columns span more than 0xFFFE (the int32 columns must then be
// used).  One wavefront per slice.
constexpr int kPad16 = 0xFFFF;

__global__ __launch_bounds__(256) void sell_col16_kernel(int n, int nslices,
                                                         const int* __restrict__ off,
                                                         const int* __restrict__ rows,
                                                         const int* __restrict__ slice_ptr,
                                                         const int* __restrict__ slice_len,
                                                         const int* __restrict__ scol,
                                                         unsigned short* __restrict__ out,
                                                         int* __restrict__ slice_base,
                                                         int* overflow) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nslices * kSlice) return;  // whole wavefronts: kSlice == kWave
    const int s = (int)(idx / kSlice);
    const long base = (long)slice_ptr[s] * kSlice + (idx % kSlice);
    const int len = slice_len[s];
    const long row = idx < n ? (rows != nullptr ? rows[idx] : idx) : -1;
    const int own = row >= 0 ? off[row + 1] - off[row] : 0;
    int lo = INT_MAX, hi = INT_MIN;
    for (int k = 0; k < own && k < len; ++k) {
        const int c = scol[base + (long)k * kSlice];
        lo = min(lo, c);
        hi = max(hi, c);
    }
    for (int m = kWave / 2; m > 0; m >>= 1) {
        lo = min(lo, __shfl_xor(lo, m, kWave));
        hi = max(hi, __shfl_xor(hi, m, kWave));
    }
    const int b0 = lo == INT_MAX ? 0 : lo;
    for (int k = 0; k < len; ++k) {
        const long p = base + (long)k * kSlice;
        out[p] = (unsigned short)(k < own ? scol[p] - b0 : kPad16);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        slice_base[s] = b0;
        if (lo != INT_MAX && (long)hi - lo > kPad16 - 1) atomicAdd(overflow, 1);
    }
}

// Stencil form of the slices (natural row order): the offsets d = column - row of
// a slice's rows, merged in ascending order, give the slice's entries; entry k of
// every row of slice s has column row + deltas[32 s + k] where bit k of
// the row's mask is set, and is padding (value 0, no gather) where it is clear.
// Entries then sit in the order of their columns, as in the plain slices.  One
// wavefront per slice walks the union with wavefront minima (each row's columns
// ascend, so each step advances at least one row).  Without slice_ptr it only
// counts (slice_len[s] = union size; status[0] += slices with more than 32
// offsets, which a 32-bit mask cannot hold); with slice_ptr it writes deltas, the
// row masks and the map (every position of the slice's union, -1 for padding).
constexpr int kStencilMax = 32;

__global__ __launch_bounds__(256) void sell_stencil_kernel(
    int n, int nslices, const int* __restrict__ off, const int* __restrict__ col,
    const int* __restrict__ slice_ptr, int* __restrict__ slice_len, int* __restrict__ deltas,
    int* __restrict__ row_mask, int* __restrict__ smap, int* status) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nslices * kSlice) return;  // whole wavefronts: kSlice == kWave
    const int s = (int)(idx / kSlice);
    const int lane = (int)(idx % kSlice);
    const bool fill = slice_ptr != nullptr;
    const long sp = fill ? (long)slice_ptr[s] : 0;
    const int row = (int)idx;
    int j = row < n ? off[row] : 0;
    const int b = row < n ? off[row + 1] : 0;
    unsigned mask = 0;
    int count = 0;
    for (;;) {
        const int d = j < b ? col[j] - row : INT_MAX;
        int m = d;
        for (int w = kWave / 2; w > 0; w >>= 1) m = min(m, __shfl_xor(m, w, kWave));
        if (m == INT_MAX) break;  // (uniform: every lane holds the minimum)
        const bool has = d == m;
        if (fill && count < kStencilMax) {
            smap[(sp + count) * kSlice + lane] = has ? j : -1;
            if (lane == 0) deltas[(long)s * kStencilMax + count] = m;
        }
        if (has) {
            if (count < kStencilMax) mask |= 1u << count;
            ++j;
        }
        ++count;
    }
    if (!fill) {
        if (lane == 0) {
            slice_len[s] = count;
            if (count > kStencilMax) atomicAdd(status, 1);
        }
    } else if (row < n) {
        row_mask[row] = (int)mask;
    }
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

int wpc_sell_col16(int n, const int* offsets, const int* rows, const int* slice_ptr,
                   const int* slice_len, const int* sell_columns, short* sell_col16,
                   int* slice_base, int* overflow, void* stream) {
    const int ns = wpc_sell_slices(n);
    if (ns == 0) return WPC_OK;
    hipLaunchKernelGGL(sell_col16_kernel, dim3(blocks_for((long)ns * kSlice, 256)), dim3(256), 0,
                       as_stream(stream), n, ns, offsets, rows, slice_ptr, slice_len,
                       sell_columns, (unsigned short*)sell_col16, slice_base, overflow);
    return launch_status();
}

int wpc_sell_stencil(int n, const int* offsets, const int* columns, const int* slice_ptr,
                     int* slice_len, int* deltas, int* row_mask, int* sell_map, int* status,
                     void* stream) {
    const int ns = wpc_sell_slices(n);
    if (ns == 0) return WPC_OK;
    if (offsets == nullptr || columns == nullptr || slice_len == nullptr || status == nullptr)
        return WPC_EARG;
    if (slice_ptr != nullptr && (deltas == nullptr || row_mask == nullptr || sell_map == nullptr))
        return WPC_EARG;
    hipLaunchKernelGGL(sell_stencil_kernel, dim3(blocks_for((long)ns * kSlice, 256)), dim3(256), 0,
                       as_stream(stream), n, ns, offsets, columns, slice_ptr, slice_len, deltas,
                       row_mask, sell_map, status);
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
