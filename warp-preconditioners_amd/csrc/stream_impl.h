// Streamed CSR products ("CSR-stream"): the hot sparse kernel of every apply.
//
// Rows are pre-partitioned into blocks of at most 256 rows and kStreamNnz stored
// entries.  A workgroup loads its block's contiguous column/value range into LDS
// with fully coalesced loads (the HBM stream), then one thread per row sums its
// entries serially in storage order while gathering x (nvec interleaved vectors,
// mostly cache hits).  A row longer than kStreamNnz gets a block of its own and
// is reduced by the whole workgroup.
//
// Epilogue:  z_i = alpha * (s_i * sum_i + d_i * u_i) + beta * y_i   (s, d, u optional)
// covers the FSAI passes (s = per-row scale^2 of a narrow factor), the factored
// operator's two passes (s = 1/mass; d = w*mass, u = x) and plain products.  An
// optional fused inner product writes one partial per (block, vector); a
// single-workgroup launch folds them in block order (deterministic).


// Layout of the code: this header holds the kernels and their launch templates;
// each stream_nv<N>.hip instantiates them for N interleaved vectors (so the eight
// instantiation sets compile in parallel), and stream.hip holds the C entry points.

#pragma once

#include "reduce.h"

namespace wpc {

constexpr int kStreamNnz = 2048;

struct StreamArgs {
    int nb;
    const int* blocks;  // streamed mapping: row partition; sliced mapping: slice_ptr
    const int* slice_len;  // sliced mapping only
    const int* slice_base; // sliced mapping, 16-bit columns: smallest column per slice
                           // (stencil deltas: the row masks)
    const int* rows;       // sliced mapping: row order of the slices (NULL: natural)
    int col16;             // sliced mapping: column form (kColI32, kColU16, kColStencil)
    int n;
    int lg;  // log2(lanes per row), -1 for the streamed mapping, -2 for the sliced one
    const int* off;
    const int* col;
    const void* val;
    int val_dtype;
    const void* s;
    bool s32;  // s stored as float32 (else the vector type)
    const void* d;
    const void* u;
    const void* x;
    const void* y;
    void* z;
    int vec_dtype;
    double alpha;
    double beta;
    const void* w;
    double* partials;
    double* out;
    hipStream_t st;
};

int stream_entry_1(const StreamArgs& a);
int stream_entry_2(const StreamArgs& a);
int stream_entry_3(const StreamArgs& a);
int stream_entry_4(const StreamArgs& a);
int stream_entry_5(const StreamArgs& a);
int stream_entry_6(const StreamArgs& a);
int stream_entry_7(const StreamArgs& a);
int stream_entry_8(const StreamArgs& a);

// One product of a grouped launch (all fields 64-bit, in this order; the host
// builds the table as int64 words): the sliced copy, the epilogue operands, the
// vectors, the partials and the task's row count and first global workgroup.
struct SellTask {
    const int* slice_ptr;
    const int* slice_len;
    const int* rows;
    const void* col;
    const void* val;
    const void* s;
    const void* d;
    const void* u;
    const void* x;
    void* z;
    const void* w;
    double* partials;
    long n;
    long block0;
    const int* slice_base;
    long nv;  // interleaved vectors of this system (1..8) | 256 if the columns are
              // 16-bit offsets from slice_base (else int32)
};
static_assert(sizeof(SellTask) == 16 * 8, "SellTask is 16 64-bit words");

struct GroupedArgs {
    int ntasks;
    const SellTask* tasks;
    int total_blocks;
    int val_dtype;
    bool hs;
    bool s32;
    bool hd;
    bool dot;
    int vec_dtype;
    double alpha;
    hipStream_t st;
};

// Batched products, each task with its own vector count (stream_grouped.hip).
int stream_grouped_entry(const GroupedArgs& a);

// The factored operator's two sliced passes inside CG with the residual update
// fused into the second pass (wpc_sqlap_cg_update).
struct SqlapCgArgs {
    int n;
    const int* slice_ptr;
    const int* slice_len;
    const int* slice_base;
    const int* rows;
    int col16;  // column form (kColI32, kColU16, kColStencil)
    const void* col;
    const void* val;
    int val_dtype;
    const void* s;   // 1/mass, vector type
    const void* wm;  // w * mass or NULL
    const void* p;
    void* t;
    void* r;
    int vec_dtype;
    double* part_e;
    double* part_rr;
    double* state;
    int passes;  // bit 0: the first pass, bit 1: the second, bit 2: ALPHA (and with
                 // bit 3 BETA) from the state, bit 3: the first pass in recurrence form
                 // (gathers z, updates p and t)
    void* x;     // advanced by the second pass (required with bit 2)
    const void* z;          // bit 3: the preconditioned residual
    int P_rz;               // bit 3: count of the preconditioner's r.z partials
    const double* part_rz;  // bit 3: those partials
    hipStream_t st;
};

// Both passes run workgroups of kSqlapCgRows threads (eight wavefronts), which
// leave one partial per workgroup and vector: half the partials of 256-thread
// workgroups, for the second pass to fold in every one of its workgroups (those
// L2 reads compete with its gathers) and for the direction step to fold after it
// (1,024 threads made the first pass slower).
constexpr int kSqlapCgRows = 512;

int sqlap_cg_entry_1(const SqlapCgArgs& a);
int sqlap_cg_entry_2(const SqlapCgArgs& a);
int sqlap_cg_entry_3(const SqlapCgArgs& a);
int sqlap_cg_entry_4(const SqlapCgArgs& a);
int sqlap_cg_entry_5(const SqlapCgArgs& a);
int sqlap_cg_entry_6(const SqlapCgArgs& a);
int sqlap_cg_entry_7(const SqlapCgArgs& a);
int sqlap_cg_entry_8(const SqlapCgArgs& a);

// Column forms of the sliced copy: int32 columns, 16-bit offsets from a per-slice
// base, or stencil deltas (32 per slice) with per-row masks.
constexpr int kColI32 = 0, kColU16 = 1, kColStencil = 2;
constexpr int kStencilStride = 32;
// Column forms a kernel instantiation handles (each its own gather loop, so a
// kernel carries only the loops it can run): the plain forms (int32 or 16-bit,
// chosen at run time), the stencil form alone, or all three (batched launches
// whose systems may differ).
constexpr int kCfPlain = 0, kCfStencil = 1, kCfAny = 2;

}  // namespace wpc

#ifdef WPC_STREAM_NV

namespace wpc {
namespace {


template <class TV> struct LdsOf { using type = TV; };
template <> struct LdsOf<f16s> { using type = uint16_t; };
template <> struct LdsOf<bf16s> { using type = uint16_t; };

template <class TA, class TV> __device__ __forceinline__ TA lds_val(typename LdsOf<TV>::type v);
template <> __device__ __forceinline__ double lds_val<double, double>(double v) { return v; }
template <> __device__ __forceinline__ double lds_val<double, float>(float v) { return v; }
template <> __device__ __forceinline__ float lds_val<float, float>(float v) { return v; }
template <> __device__ __forceinline__ double lds_val<double, f16s>(uint16_t v) { return f16_to_float(v); }
template <> __device__ __forceinline__ float lds_val<float, f16s>(uint16_t v) { return f16_to_float(v); }
template <> __device__ __forceinline__ double lds_val<double, bf16s>(uint16_t v) { return bf16_to_float(v); }
template <> __device__ __forceinline__ float lds_val<float, bf16s>(uint16_t v) { return bf16_to_float(v); }

template <class TV> __device__ __forceinline__ typename LdsOf<TV>::type raw_val(const TV* p, long i);
template <> __device__ __forceinline__ double raw_val<double>(const double* p, long i) { return p[i]; }
template <> __device__ __forceinline__ float raw_val<float>(const float* p, long i) { return p[i]; }
template <> __device__ __forceinline__ uint16_t raw_val<f16s>(const f16s* p, long i) { return p[i].b; }
template <> __device__ __forceinline__ uint16_t raw_val<bf16s>(const bf16s* p, long i) { return p[i].b; }

template <class TX> struct AccOf { using type = TX; };
template <> struct AccOf<f16s> { using type = float; };

template <class TV, class TX, class TS, int NV, bool HS, bool HD, bool DOT>
__global__ __launch_bounds__(kStreamRows) void stream_kernel(
    const int* __restrict__ blocks, const int* __restrict__ off, const int* __restrict__ col,
    const TV* __restrict__ val, const TS* __restrict__ s, const TX* __restrict__ d,
    const TX* __restrict__ u, const TX* __restrict__ x, const TX* y, TX* z,
    typename AccOf<TX>::type alpha, typename AccOf<TX>::type beta, const TX* __restrict__ w,
    double* __restrict__ partials, unsigned* counter, double* out) {
    using TA = typename AccOf<TX>::type;
    using TL = typename LdsOf<TV>::type;
    __shared__ int s_col[kStreamNnz];
    __shared__ TL s_val[kStreamNnz];
    __shared__ double red[kStreamRows];
    const int blk = xcd_block();
    const int r0 = blocks[blk], r1 = blocks[blk + 1];
    const int e0 = off[r0];
    const int nnz = off[r1] - e0;
    TA acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = TA(0);
    const bool use_a = alpha != TA(0);
    int row = r0 + threadIdx.x;
    bool own = row < r1;
    if (nnz <= kStreamNnz) {
        if (use_a) {
            for (int k = threadIdx.x; k < nnz; k += kStreamRows) {
                s_col[k] = col[e0 + k];
                s_val[k] = raw_val<TV>(val, e0 + k);
            }
        }
        __syncthreads();
        if (own && use_a) {
            const int a = off[row] - e0, b = off[row + 1] - e0;
            for (int k = a; k < b; ++k) {
                const long c = s_col[k];
                const TA av = lds_val<TA, TV>(s_val[k]);
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] += av * St<TX>::template ld<TA>(x, c * NV + v);
            }
        }
    } else {
        // one long row (r1 == r0 + 1): strided partial sums, fixed-order block reduction
        row = r0;
        own = threadIdx.x == 0;
        double part[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) part[v] = 0.0;
        if (use_a) {
            for (int k = threadIdx.x; k < nnz; k += kStreamRows) {
                const long c = col[e0 + k];
                const TA av = lds_val<TA, TV>(raw_val<TV>(val, e0 + k));
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    part[v] += (double)(av * St<TX>::template ld<TA>(x, c * NV + v));
            }
        }
        block_sum<NV>(part, red);
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = (TA)part[v];
        __syncthreads();
    }
    double dot[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) dot[v] = 0.0;
    if (own) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long k = (long)row * NV + v;
            TA r = TA(0);
            if (use_a) {
                r = acc[v];
                if (HS) r = (TA)s[row] * r;
                if (HD) r += St<TX>::template ld<TA>(d, row) * St<TX>::template ld<TA>(u, k);
                r *= alpha;
            }
            if (beta != TA(0)) r += beta * St<TX>::template ld<TA>(y, k);
            // the dot from the value as stored, and w read, before the store: a load
            // after it would make the wait for w a wait for the store as well
            TX out[1];
            St<TX>::template st<TA>(out, 0, r);
            if (DOT) dot[v] = (double)St<TX>::template ld<TA>(w, k) *
                              (double)St<TX>::template ld<TA>(out, 0);
            z[k] = out[0];
        }
    }
    if (DOT) {
        block_sum<NV>(dot, red);
        write_partials<NV>(dot, partials, blk);
    }
}

// Multi-lane mapping of the same product: L = 1 << lg lanes own one row, stride
// over its entries and reduce with an xor butterfly (DPP quad permutations for
// strides 2 and 1, shuffles above; no LDS staging).  Same epilogue and fused inner
// product as stream_kernel.
template <class TV, class TX, class TS, int NV, bool HS, bool HD, bool DOT>
__global__ __launch_bounds__(kStreamRows) void lanes_kernel(
    int n, int lg, const int* __restrict__ off, const int* __restrict__ col,
    const TV* __restrict__ val, const TS* __restrict__ s, const TX* __restrict__ d,
    const TX* __restrict__ u, const TX* __restrict__ x, const TX* y, TX* z,
    typename AccOf<TX>::type alpha, typename AccOf<TX>::type beta, const TX* __restrict__ w,
    double* __restrict__ partials) {
    using TA = typename AccOf<TX>::type;
    __shared__ double red[kStreamRows];
    const int L = 1 << lg;
    const int blk = xcd_block();
    const long row = ((long)blk * kStreamRows + threadIdx.x) >> lg;
    const int sub = threadIdx.x & (L - 1);
    const bool live = row < n;
    const bool use_a = alpha != TA(0);
    // The row's extent and the epilogue operands are loaded together, without
    // branches (lanes past n read row n-1's and discard them).
    const long rc = live ? row : n - 1;
    const int o0 = off[rc], o1 = off[rc + 1];
    const int e0 = o0, e1 = live && use_a ? o1 : o0;
    TA sv = TA(1), dv = TA(0), uv[NV], wv[NV];
    if (HS) sv = (TA)s[rc];
    if (HD) dv = St<TX>::template ld<TA>(d, rc);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if (HD) uv[v] = St<TX>::template ld<TA>(u, rc * NV + v);
        if (DOT) wv[v] = St<TX>::template ld<TA>(w, rc * NV + v);
    }
    TA acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = TA(0);
    // Two of the lane's entries per round trip (both index/value loads, then both
    // gathers); each lane still sums its entries e0+sub, e0+sub+L, ... in order.
    int e = e0 + sub;
    for (; e + L < e1; e += 2 * L) {
        const long c0 = col[e], c1 = col[e + L];
        const TA a0 = lds_val<TA, TV>(raw_val<TV>(val, e));
        const TA a1 = lds_val<TA, TV>(raw_val<TV>(val, e + L));
        TA g0[NV], g1[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            g0[v] = St<TX>::template ld<TA>(x, c0 * NV + v);
            g1[v] = St<TX>::template ld<TA>(x, c1 * NV + v);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += a0 * g0[v];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += a1 * g1[v];
    }
    if (e < e1) {
        const long c = col[e];
        const TA av = lds_val<TA, TV>(raw_val<TV>(val, e));
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += av * St<TX>::template ld<TA>(x, c * NV + v);
    }
    // Lane reduction, xor butterfly from the widest stride down; strides 2 and 1
    // stay inside a quad and run as DPP quad permutations (the same pairs).
    for (int m = L >> 1; m > 2; m >>= 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += shfl_xor(acc[v], m, kWave);
    }
    if (L >= 4) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += quad_xor<2>(acc[v]);
    }
    if (L >= 2) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += quad_xor<1>(acc[v]);
    }
    // As in the sliced mapping: every lane computes (on row rc) and only the stores
    // are predicated, so no pending load reaches a branch join ahead of the
    // reduction (where the compiler would also wait for the stores).
    const bool owner = live && sub == 0;
    double dot[NV];
    TX out[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const long k = rc * NV + v;
        TA r = TA(0);
        if (use_a) {
            r = acc[v];
            if (HS) r = sv * r;
            if (HD) r += dv * uv[v];
            r *= alpha;
        }
        if (beta != TA(0)) r += beta * St<TX>::template ld<TA>(y, k);
        St<TX>::template st<TA>(out, v, r);
        dot[v] = 0.0;
        if (DOT) {
            if (owner) dot[v] = (double)wv[v] * (double)St<TX>::template ld<TA>(out, v);
            asm volatile("" : "+v"(dot[v]));
        }
    }
    if (owner) {
#pragma unroll
        for (int v = 0; v < NV; ++v) z[row * NV + v] = out[v];
    }
    if (DOT) {
        block_sum<NV>(dot, red);
        write_partials<NV>(dot, partials, blk);
    }
}

// Sliced-ELL mapping of the same product: rows in slices of 64 (one wavefront);
// a slice's entries are stored column-major -- entry k of the row in position r of
// the slice sits at (slice_ptr[s] + k) * 64 + r -- and every row is padded to the
// slice's longest row with value 0.  Position i holds row rows[i] (rows == NULL:
// row i).  One thread per row reads fully coalesced index/value streams, issues
// its gathers a group of entries at a time, and sums in storage order, so results
// equal the streamed mapping's bit for bit.  Same epilogue and fused inner product.
//
// Gathers go through a buffer descriptor over x (32-bit byte offsets, no 64-bit
// address arithmetic).  Columns are int32 (descriptor over all of x, padding -1:
// offset >= 2^31) or unsigned 16-bit offsets from the slice's smallest column
// slice_base[s] (descriptor over 0xFFFF columns from there, padding 0xFFFF).  A
// padding offset lies outside the descriptor's range, so its load returns 0 with
// no memory access and padding adds exactly +0, whatever x holds.

// Stencil columns (column form 2): entry kk of the row has column row + delta[kk]
// (the slice's deltas, uniform across the wavefront: scalar loads) where bit kk of
// the row's mask is set; a clear bit is padding, whose offset 2^32 - stride (also
// with the vector offsets added) lies outside the descriptor over all of x.
struct StencilCol {
    const int* delta;
    unsigned mask;
    int row;
};

// Byte offset of entry kk (stored at p) of a row inside the descriptor (row
// stride `stride`).
template <class TI>
__device__ __forceinline__ unsigned sell_off(const TI* col, long p, int kk, unsigned stride);
template <>
__device__ __forceinline__ unsigned sell_off<int>(const int* col, long p, int, unsigned stride) {
    return (unsigned)col[p] * stride;
}
template <>
__device__ __forceinline__ unsigned sell_off<unsigned short>(const unsigned short* col, long p,
                                                             int, unsigned stride) {
    return (unsigned)col[p] * stride;
}


// NV > 0: NV interleaved vectors; NV == 0: nv_rt (<= 8) vectors, the count a
// batched launch reads per system.  Both perform the same operations per vector.
template <int NV> constexpr int kNvMax = NV > 0 ? NV : 8;

// Entries of a row handled per round trip: a group's index and value loads go out
// together, then its gathers.  A slice's length (uniform across the wavefront) is
// covered by groups of 8 until at most 12 entries remain; a remainder of 9-12 is
// one masked group of 12 (where 12 entries' gathers fit in registers), then 8 a
// group of 8, 5-7 a masked group of 8, 3 a masked group of 4, otherwise groups of
// 4, 2 and 1 (with many vectors, groups of 4 and the remainder 3 masked, 2 or 1:
// kSellMain).  A masked group's entries past the slice re-read its last entry (a
// cache hit; nothing past the slice is read) and their products are discarded by
// a select, so the sums are those of storage order.  Masking only where it wastes
// at most three gathers keeps the gathers (which bound these passes) close to the
// stored entries; a remainder of 5 as one masked group (not 4 + 1) saves a round
// trip, which the slowest wavefronts of a launch (every wavefront of a pass is
// resident at once) set the pass's time by.
template <class TI, class TV, class TX, int NV, int G, bool FULL = true>
__device__ __forceinline__ void sell_group(typename AccOf<TX>::type* acc, long base, int k,
                                           __amdgpu_buffer_rsrc_t xr, unsigned stride,
                                           const TI* __restrict__ col,
                                           const TV* __restrict__ val, int nv, int m = G) {
    using TA = typename AccOf<TX>::type;
    constexpr int NVM = kNvMax<NV>;
    unsigned o[G];
    TA a[G];
    if constexpr (std::is_same_v<TI, StencilCol>) {
        // The group's deltas are loaded first and pinned in scalar registers:
        // left to the compiler, each delta load sank into a branch on the row's
        // mask bit and was waited for there, one scalar round trip per entry.
        int dl[G];
#pragma unroll
        for (int j = 0; j < G; ++j)  // (uniform; in batched launches the compiler cannot tell)
            dl[j] = __builtin_amdgcn_readfirstlane(col->delta[k + (FULL || j < m ? j : m - 1)]);
#pragma unroll
        for (int j = 0; j < G; ++j) asm volatile("" : "+s"(dl[j]));
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int e = FULL || j < m ? j : m - 1;
            const long p = base + (long)(k + e) * 64;
            const unsigned on = (unsigned)(col->row + dl[j]) * stride;
            o[j] = (col->mask >> (k + e)) & 1u ? on : 0u - stride;
            a[j] = lds_val<TA, TV>(raw_val<TV>(val, p));
        }
    } else {
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const int e = FULL || j < m ? j : m - 1;
            const long p = base + (long)(k + e) * 64;
            o[j] = sell_off<TI>(col, p, k + e, stride);
            a[j] = lds_val<TA, TV>(raw_val<TV>(val, p));
        }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
#pragma unroll
        for (int v = 0; v < NVM; ++v) {
            if (v < nv) {
                const TA t = acc[v] + a[j] * (TA)buffer_ld<TX>(xr, o[j] + v * (unsigned)sizeof(TX));
                acc[v] = FULL || j < m ? t : acc[v];
            }
        }
    }
}

// Entries per full group: 8 while a group's gathered values fit in 64 registers
// (up to 4 double or 8 float vectors), else 4 (8 entries of 8 double vectors take
// 130 registers, too many for four wavefronts per SIMD).
template <class TX, int NV>
constexpr int kSellMain = kNvMax<NV> * sizeof(typename AccOf<TX>::type) <= 32 ? 8 : 4;

template <class TI, class TV, class TX, int NV>
__device__ __forceinline__ void sell_accum(typename AccOf<TX>::type* acc, long base, int len,
                                           __amdgpu_buffer_rsrc_t xr, unsigned stride,
                                           const TI* __restrict__ col,
                                           const TV* __restrict__ val, int nv_rt) {
    const int nv = NV > 0 ? NV : nv_rt;
    int k = 0;
    if constexpr (kSellMain<TX, NV> == 4) {
        for (; k + 4 <= len; k += 4)
            sell_group<TI, TV, TX, NV, 4>(acc, base, k, xr, stride, col, val, nv);
        const int r = len - k;
        if (r == 3)
            sell_group<TI, TV, TX, NV, 4, false>(acc, base, k, xr, stride, col, val, nv, 3);
        else if (r == 2)
            sell_group<TI, TV, TX, NV, 2>(acc, base, k, xr, stride, col, val, nv);
        else if (r == 1)
            sell_group<TI, TV, TX, NV, 1>(acc, base, k, xr, stride, col, val, nv);
        return;
    }
    // 12 entries' gathers: at most 72 registers (3 double or 6 float vectors)
    constexpr bool kTwelve = kNvMax<NV> * sizeof(typename AccOf<TX>::type) <= 24;
    for (; len - k > (kTwelve ? 12 : 7); k += 8)
        sell_group<TI, TV, TX, NV, 8>(acc, base, k, xr, stride, col, val, nv);
    int r = len - k;
    if (kTwelve && r > 8) {
        sell_group<TI, TV, TX, NV, 12, false>(acc, base, k, xr, stride, col, val, nv, r);
        return;
    }
    if (r == 8) {
        sell_group<TI, TV, TX, NV, 8>(acc, base, k, xr, stride, col, val, nv);
        return;
    }
    if (r >= 5) {
        sell_group<TI, TV, TX, NV, 8, false>(acc, base, k, xr, stride, col, val, nv, r);
        return;
    }
    if (r >= 4) {
        sell_group<TI, TV, TX, NV, 4>(acc, base, k, xr, stride, col, val, nv);
        k += 4;
        r -= 4;
    }
    if (r == 3) {
        sell_group<TI, TV, TX, NV, 4, false>(acc, base, k, xr, stride, col, val, nv, 3);
        return;
    }
    if (r >= 2) {
        sell_group<TI, TV, TX, NV, 2>(acc, base, k, xr, stride, col, val, nv);
        k += 2;
        r -= 2;
    }
    if (r == 1) sell_group<TI, TV, TX, NV, 1>(acc, base, k, xr, stride, col, val, nv);
}

// sell_accum in column form cm (slice: the wavefront's slice; mask, row: this
// lane's stencil mask and row, read only in the stencil form).
template <class TV, class TX, int NV, int CF>
__device__ __forceinline__ void sell_accum_cm(typename AccOf<TX>::type* acc, long base, int len,
                                              __amdgpu_buffer_rsrc_t xr, unsigned stride, int cm,
                                              const void* __restrict__ col, int slice,
                                              unsigned mask, int row, const TV* __restrict__ val,
                                              int nv) {
    if (CF == kCfStencil || (CF == kCfAny && cm == kColStencil)) {
        if constexpr (CF != kCfPlain) {
            const StencilCol sc{(const int*)col + (long)slice * kStencilStride, mask, row};
            sell_accum<StencilCol, TV, TX, NV>(acc, base, len, xr, stride, &sc, val, nv);
        }
    } else if (cm == kColU16) {
        sell_accum<unsigned short, TV, TX, NV>(acc, base, len, xr, stride,
                                               (const unsigned short*)col, val, nv);
    } else {
        sell_accum<int, TV, TX, NV>(acc, base, len, xr, stride, (const int*)col, val, nv);
    }
}

// Descriptor range over x (from the slice base) for column form cm.
__device__ __forceinline__ int sell_range(int cm, unsigned stride) {
    return (int)(cm == kColU16 ? 0xFFFFu * stride : kRange32);
}

// blk must be uniform across the workgroup (the slice header is read with scalar
// loads, and the gather descriptor is built from it).  cm: the column form
// (kColU16: 16-bit offsets from slice_base[s]; kColI32: int32 columns and
// slice_base holds zeros; kColStencil: deltas, slice_base holds the row masks,
// natural row order).
template <class TV, class TX, class TS, int NV, bool HS, bool HD, bool DOT, int CF>
__device__ __forceinline__ void sell_body(
    int blk, int n, const int* __restrict__ slice_ptr, const int* __restrict__ slice_len,
    const int* __restrict__ slice_base, const int* __restrict__ rows,
    const void* __restrict__ col, int cm,
    const TV* __restrict__ val, const TS* __restrict__ s, const TX* __restrict__ d,
    const TX* __restrict__ u, const TX* __restrict__ x, const TX* y, TX* z,
    typename AccOf<TX>::type alpha, typename AccOf<TX>::type beta, const TX* __restrict__ w,
    double* __restrict__ partials, double* red, int nv_rt = NV) {
    using TA = typename AccOf<TX>::type;
    constexpr int NVM = kNvMax<NV>;
    const int nv = NV > 0 ? NV : nv_rt;
    const int idx = blk * kStreamRows + threadIdx.x;
    const bool live = idx < n;
    // The wavefront's slice (uniform: a slice is one wavefront; a wavefront past
    // the last row works on the last slice and stores nothing).  Its three header
    // words are scalar loads issued together, with no condition between them and
    // the kernel arguments.  Lanes past n read row n-1's epilogue operands and
    // discard them; the padding positions of the last slice they gather from read 0.
    const int wave0 = __builtin_amdgcn_readfirstlane(idx & ~63);
    const int slice = min(wave0, n - 1) >> 6;
    const int sp = slice_ptr[slice];
    const int sl = slice_len[slice];
    const bool stencil = CF == kCfStencil || (CF == kCfAny && cm == kColStencil);
    const int sb = stencil ? 0 : slice_base[slice];
    const int idc = live ? idx : n - 1;
    const unsigned mask = stencil ? (unsigned)slice_base[idc] : 0u;
    const int row = rows != nullptr ? rows[idc] : idc;
    const bool use_a = alpha != TA(0);
    const long base = (long)sp * 64 + (idx & 63);
    const int len = use_a && wave0 < n ? sl : 0;
    const unsigned stride = (unsigned)(nv * sizeof(TX));
    // (slice_base holds zeros with int32 columns, so the base needs no branch)
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(x + (long)sb * nv), (short)0, sell_range(cm, stride), kBufferWord3);
    TA sv = TA(1), dv = TA(0), uv[NVM], wv[NVM];
    if (HS) sv = (TA)s[row];
    if (HD) dv = St<TX>::template ld<TA>(d, row);
#pragma unroll
    for (int v = 0; v < NVM; ++v) {
        if (v < nv) {
            if (HD) uv[v] = St<TX>::template ld<TA>(u, (long)row * nv + v);
            if (DOT) wv[v] = St<TX>::template ld<TA>(w, (long)row * nv + v);
        }
    }
    TA acc[NVM];
#pragma unroll
    for (int v = 0; v < NVM; ++v) acc[v] = TA(0);
    // (len is uniform: the entry loop's control is scalar)
    sell_accum_cm<TV, TX, NV, CF>(acc, base, len, xr, stride, cm, col, slice, mask, row, val, nv);
    // Every lane computes (a lane past n on row n-1's operands) and only the stores
    // are predicated: a branch around the loads would leave them pending at its
    // join, where the compiler then waits for every outstanding memory operation
    // (these stores included) before the reduction below.
    double dot[NVM];
    TX out[NVM];
#pragma unroll
    for (int v = 0; v < NVM; ++v) {
        dot[v] = 0.0;
        if (v >= nv) break;
        const long k = (long)row * nv + v;
        TA r = TA(0);
        if (use_a) {
            r = acc[v];
            if (HS) r = sv * r;
            if (HD) r += dv * uv[v];
            r *= alpha;
        }
        if (beta != TA(0)) r += beta * St<TX>::template ld<TA>(y, k);
        St<TX>::template st<TA>(out, v, r);
        // (the dot takes z as stored; the empty asm keeps the product, and the wait
        // for w, ahead of the stores instead of sunk past them)
        if (DOT) {
            if (live) dot[v] = (double)wv[v] * (double)St<TX>::template ld<TA>(out, v);
            asm volatile("" : "+v"(dot[v]));
        }
    }
    if (live) {
#pragma unroll
        for (int v = 0; v < NVM; ++v) {
            if (v >= nv) break;
            z[(long)row * nv + v] = out[v];
        }
    }
    if (DOT) {
        if constexpr (NV > 0) {
            block_sum<NV>(dot, red);
            write_partials<NV>(dot, partials, blk);
        } else {
            block_sum_n(dot, nv, red);
            write_partials_n(dot, nv, partials, blk);
        }
    }
}

template <class TV, class TX, class TS, int NV, bool HS, bool HD, bool DOT, int CF>
__global__ __launch_bounds__(kStreamRows) void sell_kernel(
    int n, const int* __restrict__ slice_ptr, const int* __restrict__ slice_len,
    const int* __restrict__ slice_base, const int* __restrict__ rows,
    const void* __restrict__ col, int col16,
    const TV* __restrict__ val, const TS* __restrict__ s, const TX* __restrict__ d,
    const TX* __restrict__ u, const TX* __restrict__ x, const TX* y, TX* z,
    typename AccOf<TX>::type alpha, typename AccOf<TX>::type beta, const TX* __restrict__ w,
    double* __restrict__ partials) {
    __shared__ double red[kStreamRows];
    sell_body<TV, TX, TS, NV, HS, HD, DOT, CF>(xcd_block(), n, slice_ptr, slice_len, slice_base, rows,
                                           col, col16, val, s, d, u, x, y, z, alpha, beta, w,
                                           partials, red);
}

// The factored operator A = L M^-1 L + W (W = w M, diagonal; L symmetric) inside
// CG, with the residual update fused into its second pass.  With t = M^-1 L p the
// curvature is
//     p.Ap = sum_i (L p)_i t_i + W_i p_i^2,
// so the first pass leaves partials of it and the second pass, which folds them
// (fixed order, every workgroup alike, as the update step folds p.Ap), knows
// ALPHA before its epilogue: it forms Ap_i = (L t)_i + W_i p_i exactly as the plain
// second pass does and applies r -= ALPHA Ap instead of storing Ap, which no later
// step reads.  The update launch and Ap's round trip through HBM leave the
// iteration.  State fields as in the update step: reads RZ, RR, TARGET; workgroup 0
// writes RZ_OLD (= RZ), DENOM, ACTIVE, ALPHA; given x, the second pass also
// advances x += ALPHA p (it reads p anyway).

// Wavefront slice header and gather descriptor of the sliced mapping (as sell_body).
struct SellHead {
    int row, len, slice;
    unsigned mask;  // the row's stencil mask (column form kColStencil)
    bool live;
    long base;
};

template <class TX, int NV, int NT, int CF>
__device__ __forceinline__ SellHead sell_head(int blk, int n, const int* __restrict__ slice_ptr,
                                              const int* __restrict__ slice_len,
                                              const int* __restrict__ slice_base,
                                              const int* __restrict__ rows, int cm,
                                              const TX* x, __amdgpu_buffer_rsrc_t* xr) {
    const int idx = blk * NT + threadIdx.x;
    const int wave0 = __builtin_amdgcn_readfirstlane(idx & ~63);
    const int slice = min(wave0, n - 1) >> 6;
    const int sp = slice_ptr[slice];
    const int sl = slice_len[slice];
    const bool stencil = CF == kCfStencil || (CF == kCfAny && cm == kColStencil);
    const int sb = stencil ? 0 : slice_base[slice];
    const int idc = idx < n ? idx : n - 1;
    SellHead h;
    h.live = idx < n;
    h.row = rows != nullptr ? rows[idc] : idc;
    h.mask = stencil ? (unsigned)slice_base[idc] : 0u;
    h.slice = slice;
    h.base = (long)sp * 64 + (idx & 63);
    h.len = wave0 < n ? sl : 0;
    const unsigned stride = (unsigned)(NV * sizeof(TX));
    *xr = __builtin_amdgcn_make_buffer_rsrc((void*)(x + (long)sb * NV), (short)0,
                                            sell_range(cm, stride), kBufferWord3);
    return h;
}

// Pass 1: t = s .* (L p) (s = 1/mass), partials of (L p)_i t_i + W_i p_i^2.

template <class TV, class TX, int NV, bool HW, int CF>
__global__ __launch_bounds__(kSqlapCgRows) void sell_energy_kernel(
    int n, const int* __restrict__ slice_ptr, const int* __restrict__ slice_len,
    const int* __restrict__ slice_base, const int* __restrict__ rows,
    const void* __restrict__ col, int col16, const TV* __restrict__ val,
    const TX* __restrict__ s, const TX* __restrict__ wm, const TX* __restrict__ p,
    TX* __restrict__ t, double* __restrict__ partials) {
    using TA = typename AccOf<TX>::type;
    __shared__ double red[kStreamRows];
    const int blk = xcd_block();
    __amdgpu_buffer_rsrc_t xr;
    const SellHead h = sell_head<TX, NV, kSqlapCgRows, CF>(blk, n, slice_ptr, slice_len, slice_base,
                                                      rows, col16, p, &xr);
    const unsigned stride = (unsigned)(NV * sizeof(TX));
    const TA sv = St<TX>::template ld<TA>(s, h.row);
    TA wv = TA(0), pv[NV];
    if (HW) {
        wv = St<TX>::template ld<TA>(wm, h.row);
#pragma unroll
        for (int v = 0; v < NV; ++v) pv[v] = St<TX>::template ld<TA>(p, (long)h.row * NV + v);
    }
    TA acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = TA(0);
    sell_accum_cm<TV, TX, NV, CF>(acc, h.base, h.len, xr, stride, col16, col, h.slice, h.mask,
                                  h.row,
                              val, NV);
    double e[NV];
    TX out[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        // the plain first pass's arithmetic (alpha = 1, beta = 0)
        St<TX>::template st<TA>(out, v, sv * acc[v]);
        double ev = (double)acc[v] * (double)out[v];
        if (HW) ev += (double)wv * ((double)pv[v] * (double)pv[v]);
        e[v] = h.live ? ev : 0.0;
        asm volatile("" : "+v"(e[v]));
    }
    if (h.live) {
#pragma unroll
        for (int v = 0; v < NV; ++v) t[(long)h.row * NV + v] = out[v];
    }
    block_sum<NV, kSqlapCgRows>(e, red);
    write_partials<NV>(e, partials, blk);
}

// Pass 1 in recurrence form, with the previous step's direction step folded in:
// p' = z + BETA p and t' = M^-1 L p' = M^-1 L z + BETA t (t as the previous step's
// first pass or this kernel left it), so the pass gathers z instead of p and the
// direction launch leaves the iteration.  Every workgroup folds the r.z partials
// of the preconditioner's last pass into BETA = RZ_NEW / RZ_OLD exactly as the
// direction step does; workgroup 0 also folds the r.r partials of pass 2 and
// records RZ_NEW, RR_NEW, BETA and, for stepping slots (ACTIVE), RZ, RR and
// ITER += 1.  A resting slot keeps p and t.  Curvature partials as pass 1,
// (L p')_i t'_i + W_i p'_i^2 with (L p')_i = t'_i / s_i.  FS (for partial counts
// too large to fold in every workgroup): BETA and ACTIVE as the scalar fold
// wpc_cg_beta left them in the state (no fold, no state writes).
template <class TV, class TX, int NV, bool HW, bool FS, int CF>
__global__ __launch_bounds__(kSqlapCgRows) void sell_direction_energy_kernel(
    int n, const int* __restrict__ slice_ptr, const int* __restrict__ slice_len,
    const int* __restrict__ slice_base, const int* __restrict__ rows,
    const void* __restrict__ col, int col16, const TV* __restrict__ val,
    const TX* __restrict__ s, const TX* __restrict__ wm, const TX* __restrict__ z,
    TX* __restrict__ p, TX* __restrict__ t, int P_rz, const double* __restrict__ part_rz,
    int P_rr, const double* __restrict__ part_rr, double* state,
    double* __restrict__ partials) {
    using TA = typename AccOf<TX>::type;
    __shared__ double red[kStreamRows];
    __shared__ double s_rz[8], s_rr[8];
    const int blk = xcd_block();
    __amdgpu_buffer_rsrc_t xr;
    const SellHead h = sell_head<TX, NV, kSqlapCgRows, CF>(blk, n, slice_ptr, slice_len, slice_base,
                                                      rows, col16, z, &xr);
    const unsigned stride = (unsigned)(NV * sizeof(TX));
    const TA sv = St<TX>::template ld<TA>(s, h.row);
    TA wv = TA(0);
    if (HW) wv = St<TX>::template ld<TA>(wm, h.row);
    TA acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = TA(0);
    sell_accum_cm<TV, TX, NV, CF>(acc, h.base, h.len, xr, stride, col16, col, h.slice, h.mask,
                                  h.row,
                              val, NV);
    // After the gathers: the state fields, the partials and the row's z, p and t,
    // in one round trip.
    StateLanes<WPC_KS_ACTIVE, (FS ? WPC_KS_BETA : WPC_KS_RZ_OLD)> st;
    st.load(state, NV);
    FoldRegs<kSqlapCgRows, NV> fold_z, fold_r;
    if constexpr (!FS) {
        fold_z.load(part_rz, P_rz);
        if (blk == 0) fold_r.load(part_rr, P_rr);  // r.r is only recorded, by workgroup 0
    }
    TX zv[NV], pv[NV], tv[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        zv[v] = z[(long)h.row * NV + v];
        pv[v] = p[(long)h.row * NV + v];
        tv[v] = t[(long)h.row * NV + v];
    }
    if constexpr (!FS) {
        if (blk != 0) {
            fold_z.finish(part_rz, P_rz, s_rz);
        } else if constexpr (NV <= 4) {
            // both folds through one cross-wavefront pass
            double both_acc[8];
            __shared__ double both[8];
            fold_z.sums(part_rz, P_rz, both_acc);
            fold_r.sums(part_rr, P_rr, both_acc + NV);
            fold_reduce<kSqlapCgRows>(both_acc, 2 * NV, both);
#pragma unroll
            for (int v = 0; v < NV; ++v) { s_rz[v] = both[v]; s_rr[v] = both[NV + v]; }
        } else {
            fold_z.finish(part_rz, P_rz, s_rz);
            fold_r.finish(part_rr, P_rr, s_rr);
        }
    }
    double e[NV];
    TX pn[NV], tn[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const bool act = st.get(0, NV, v) != 0.0;
        double b;
        if constexpr (FS) {
            b = act ? st.get(1, NV, v) : 0.0;
        } else {
            b = act ? s_rz[v] / st.get(1, NV, v) : 0.0;
            if (blk == 0 && threadIdx.x == v) {
                state[WPC_KS_RZ_NEW * NV + v] = s_rz[v];
                state[WPC_KS_RR_NEW * NV + v] = s_rr[v];
                state[WPC_KS_BETA * NV + v] = b;
                if (act) {
                    state[WPC_KS_RZ * NV + v] = s_rz[v];
                    state[WPC_KS_RR * NV + v] = s_rr[v];
                    state[WPC_KS_ITER * NV + v] += 1.0;
                }
            }
        }
        const TX beta = (TX)b;
        TX tz[1];
        St<TX>::template st<TA>(tz, 0, sv * acc[v]);  // M^-1 L z as the plain pass rounds it
        pn[v] = act ? zv[v] + beta * pv[v] : pv[v];
        tn[v] = act ? tz[0] + beta * tv[v] : tv[v];
        double ev = (double)tn[v] * (double)tn[v] / (double)sv;
        if (HW) ev += (double)wv * ((double)pn[v] * (double)pn[v]);
        e[v] = h.live ? ev : 0.0;
        asm volatile("" : "+v"(e[v]));
    }
    if (h.live) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            p[(long)h.row * NV + v] = pn[v];
            t[(long)h.row * NV + v] = tn[v];
        }
    }
    block_sum<NV, kSqlapCgRows>(e, red);
    write_partials<NV>(e, partials, blk);
}

// Pass 2: Ap = L t + W p (not stored); ALPHA from the P pass-1 partials;
// r -= ALPHA Ap for stepping slots; partials of r.r.  FS (for partial counts too
// large to fold in every workgroup): ALPHA as the scalar fold wpc_cg_alpha left it
// in the state (no fold, no state writes), and x += ALPHA p here, as the update
// step would.
template <class TV, class TX, int NV, bool HW, bool FS, int CF>
__global__ __launch_bounds__(kSqlapCgRows) void sell_cg_update_kernel(
    int n, const int* __restrict__ slice_ptr, const int* __restrict__ slice_len,
    const int* __restrict__ slice_base, const int* __restrict__ rows,
    const void* __restrict__ col, int col16, const TV* __restrict__ val,
    const TX* __restrict__ wm, const TX* __restrict__ p, const TX* __restrict__ t,
    TX* __restrict__ r, int P, const double* __restrict__ part_e, double* state,
    double* __restrict__ part_rr, TX* __restrict__ x) {
    using TA = typename AccOf<TX>::type;
    __shared__ double red[kStreamRows];
    __shared__ double sums[8];
    const int blk = xcd_block();
    __amdgpu_buffer_rsrc_t xr;
    const SellHead h = sell_head<TX, NV, kSqlapCgRows, CF>(blk, n, slice_ptr, slice_len,
                                                       slice_base, rows, col16, t, &xr);
    const unsigned stride = (unsigned)(NV * sizeof(TX));
    TA wv = TA(0), pv[NV];
    if (HW) {
        wv = St<TX>::template ld<TA>(wm, h.row);
#pragma unroll
        for (int v = 0; v < NV; ++v) pv[v] = St<TX>::template ld<TA>(p, (long)h.row * NV + v);
    }
    TA acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = TA(0);
    sell_accum_cm<TV, TX, NV, CF>(acc, h.base, h.len, xr, stride, col16, col, h.slice, h.mask,
                                  h.row,
                              val, NV);
    TX alpha[NV];
    TX rv[NV], xv[NV], px[NV];
    if constexpr (FS) {
        // the residual, x and p rows and ALPHA, in one round trip after the gathers
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            rv[v] = r[(long)h.row * NV + v];
            xv[v] = x[(long)h.row * NV + v];
            px[v] = HW ? (TX)pv[v] : p[(long)h.row * NV + v];  // (TA is TX here)
            alpha[v] = (TX)state[WPC_KS_ALPHA * NV + v];
        }
    } else {
        // After the gathers (their registers are free again): the state fields, the
        // partials and the residual rows, all in one round trip.
        StateLanes<WPC_KS_RZ, WPC_KS_RR, WPC_KS_TARGET> st;
        st.load(state, NV);
        FoldRegs<kSqlapCgRows, NV> fold;
        fold.load(part_e, P);
#pragma unroll
        for (int v = 0; v < NV; ++v) rv[v] = r[(long)h.row * NV + v];
        if (x != nullptr) {  // (uniform) x += ALPHA p here as well
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                xv[v] = x[(long)h.row * NV + v];
                px[v] = HW ? (TX)pv[v] : p[(long)h.row * NV + v];
            }
        }
        fold.finish(part_e, P, sums);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double denom = sums[v];
            const double rz = st.get(0, NV, v);
            const bool active = st.get(1, NV, v) > st.get(2, NV, v);
            const double a = active ? rz / denom : 0.0;
            alpha[v] = (TX)a;
            if (blk == 0 && threadIdx.x == v) {
                state[WPC_KS_RZ_OLD * NV + v] = rz;
                state[WPC_KS_DENOM * NV + v] = denom;
                state[WPC_KS_ACTIVE * NV + v] = active ? 1.0 : 0.0;
                state[WPC_KS_ALPHA * NV + v] = a;
            }
        }
    }
    double rr[NV];
    TX rn[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        // Ap as the plain second pass stores it (alpha = 1, beta = 0), then the
        // update step's arithmetic
        TA ap = acc[v];
        if (HW) ap += wv * pv[v];
        TX apx[1];
        St<TX>::template st<TA>(apx, 0, ap);
        rn[v] = alpha[v] != TX(0) ? rv[v] - alpha[v] * apx[0] : rv[v];
        rr[v] = h.live ? (double)rn[v] * (double)rn[v] : 0.0;
    }
    if (h.live) {
#pragma unroll
        for (int v = 0; v < NV; ++v) r[(long)h.row * NV + v] = rn[v];
        if (FS || x != nullptr) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
                if (alpha[v] != TX(0)) x[(long)h.row * NV + v] = xv[v] + alpha[v] * px[v];
        }
    }
    block_sum<NV, kSqlapCgRows>(rr, red);
    write_partials<NV>(rr, part_rr, blk);
}

// The same product on several systems in one launch: workgroup g (after the XCD
// renumbering) belongs to the task with the largest block0 <= g and runs that
// task's local block g - block0, so every system computes exactly what its own
// launch would (same rows per workgroup, same partials), bit for bit.  beta = 0.
// Tasks with n == 0 pad a captured launch to its capacity and do nothing.
template <class TV, class TX, class TS, int NV, bool HS, bool HD, bool DOT>
__global__ __launch_bounds__(kStreamRows) void grouped_sell_kernel(
    const SellTask* __restrict__ tasks, int ntasks, typename AccOf<TX>::type alpha) {
    using TA = typename AccOf<TX>::type;
    __shared__ double red[kStreamRows];
    __shared__ int s_task;
    const int g = xcd_block();
    if (threadIdx.x == 0) {
        int lo = 0, hi = ntasks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (tasks[mid].block0 <= g) lo = mid;
            else hi = mid - 1;
        }
        s_task = lo;
    }
    __syncthreads();
    // (uniform: the task's fields are read with scalar loads)
    const SellTask& t = tasks[__builtin_amdgcn_readfirstlane(s_task)];
    if (t.n == 0) return;  // capacity padding of a captured launch
    const int blk = g - (int)t.block0;
    sell_body<TV, TX, TS, NV, HS, HD, DOT, kCfAny>(
        blk, (int)t.n, t.slice_ptr, t.slice_len, t.slice_base, t.rows, t.col, (t.nv >> 8) & 3,
        (const TV*)t.val, (const TS*)t.s,
        (const TX*)t.d, (const TX*)t.u, (const TX*)t.x, nullptr, (TX*)t.z, alpha, TA(0),
        (const TX*)t.w, t.partials, red, (int)(t.nv & 0xFF));
}

template <class TV, class TX, class TS, int NV>
int grouped_launch(const GroupedArgs& a) {
    using TA = typename AccOf<TX>::type;
    if (a.ntasks <= 0 || a.total_blocks <= 0) return WPC_OK;
#define GL(HS, HD, DT)                                                                          \
    hipLaunchKernelGGL((grouped_sell_kernel<TV, TX, TS, NV, HS, HD, DT>), dim3(a.total_blocks), \
                       dim3(kStreamRows), 0, a.st, a.tasks, a.ntasks, (TA)a.alpha)
    if (!a.hs && !a.hd && !a.dot) GL(false, false, false);
    else if (a.hs && !a.hd && !a.dot) GL(true, false, false);
    else if (!a.hs && a.hd && !a.dot) GL(false, true, false);
    else if (a.hs && a.hd && !a.dot) GL(true, true, false);
    else if (!a.hs && !a.hd && a.dot) GL(false, false, true);
    else if (a.hs && !a.hd && a.dot) GL(true, false, true);
    else if (!a.hs && a.hd && a.dot) GL(false, true, true);
    else GL(true, true, true);
#undef GL
    return launch_status();
}

// One vector format per translation unit (stream_grouped.hip: float64 vectors,
// stream_grouped32.hip: float32), so the two build in parallel.
template <int NV, class TX>
int grouped_for_nv(const GroupedArgs& a) {
#define GT(TV)                                                                                  \
    do {                                                                                        \
        if (a.hs && a.s32) return grouped_launch<TV, TX, float, NV>(a);                         \
        return grouped_launch<TV, TX, TX, NV>(a);                                               \
    } while (0)
    if (a.val_dtype == WPC_F32) GT(float);
    if (a.val_dtype == WPC_F16) GT(f16s);
    if (a.val_dtype == WPC_BF16) GT(bf16s);
    if constexpr (sizeof(TX) == 8) {
        if (a.val_dtype == WPC_F64) GT(double);
    }
#undef GT
    return WPC_EDTYPE;
}

// Fixed-order fold of per-block partials by one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void fold_kernel(int P, int nvec,
                                                  const double* __restrict__ partials,
                                                  double* __restrict__ out) {
    __shared__ double red[1024 / kWave][8];
    double acc[8];
    for (int v = 0; v < 8; ++v) acc[v] = 0.0;
    for (int b = threadIdx.x; b < P; b += 1024)
        for (int v = 0; v < nvec; ++v) acc[v] += partials[(long)b * nvec + v];
    for (int v = 0; v < nvec; ++v) {
        const double x = wave_sum(acc[v]);
        if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave][v] = x;
    }
    __syncthreads();
    if (threadIdx.x < nvec) {
        double t = 0.0;
        for (int w = 0; w < 1024 / kWave; ++w) t += red[w][threadIdx.x];
        out[threadIdx.x] = t;
    }
}

template <class TV, class TX, class TS, int NV>
int stream_launch(int nb, const int* blocks, const int* slice_len, const int* slice_base,
                  const int* rows, int col16, int n, int lg, const int* off,
                  const int* col, const void* val, const void* s, const void* d, const void* u,
                  const void* x, const void* y, void* z, double alpha, double beta, const void* w,
                  double* partials, double* out, hipStream_t st) {
    using TA = typename AccOf<TX>::type;
    const bool hs = s != nullptr, hd = d != nullptr, dt = w != nullptr;
    const int grid = lg == -1 ? nb : (int)blocks_for((long)n << (lg < 0 ? 0 : lg), kStreamRows);
    if (grid <= 0 || (lg == -1 && nb <= 0) || (lg != -1 && n <= 0)) return WPC_OK;
#define SL(HS, HD, DT)                                                                          \
    do {                                                                                        \
        if (lg == -2 && col16 == kColStencil)                                                   \
            hipLaunchKernelGGL((sell_kernel<TV, TX, TS, NV, HS, HD, DT, kCfStencil>), dim3(grid), \
                               dim3(kStreamRows), 0, st, n, blocks, slice_len, slice_base,      \
                               rows, (const void*)col, (int)col16,                              \
                               (const TV*)val, (const TS*)s, (const TX*)d, (const TX*)u,        \
                               (const TX*)x, (const TX*)y, (TX*)z, (TA)alpha, (TA)beta,         \
                               (const TX*)w, partials);                                         \
        else if (lg == -2)                                                                      \
            hipLaunchKernelGGL((sell_kernel<TV, TX, TS, NV, HS, HD, DT, kCfPlain>), dim3(grid), \
                               dim3(kStreamRows), 0, st, n, blocks, slice_len, slice_base,      \
                               rows, (const void*)col, (int)col16,                              \
                               (const TV*)val, (const TS*)s, (const TX*)d, (const TX*)u,        \
                               (const TX*)x, (const TX*)y, (TX*)z, (TA)alpha, (TA)beta,         \
                               (const TX*)w, partials);                                         \
        else if (lg == -1)                                                                      \
            hipLaunchKernelGGL((stream_kernel<TV, TX, TS, NV, HS, HD, DT>), dim3(nb),           \
                               dim3(kStreamRows), 0, st, blocks, off, col, (const TV*)val,      \
                               (const TS*)s, (const TX*)d, (const TX*)u, (const TX*)x,          \
                               (const TX*)y, (TX*)z, (TA)alpha, (TA)beta, (const TX*)w,         \
                               partials, nullptr, out);                                         \
        else                                                                                    \
            hipLaunchKernelGGL((lanes_kernel<TV, TX, TS, NV, HS, HD, DT>), dim3(grid),          \
                               dim3(kStreamRows), 0, st, n, lg, off, col, (const TV*)val,       \
                               (const TS*)s, (const TX*)d, (const TX*)u, (const TX*)x,          \
                               (const TX*)y, (TX*)z, (TA)alpha, (TA)beta, (const TX*)w,         \
                               partials);                                                       \
    } while (0)
    if (!hs && !hd && !dt) SL(false, false, false);
    else if (hs && !hd && !dt) SL(true, false, false);
    else if (!hs && hd && !dt) SL(false, true, false);
    else if (hs && hd && !dt) SL(true, true, false);
    else if (!hs && !hd && dt) SL(false, false, true);
    else if (hs && !hd && dt) SL(true, false, true);
    else if (!hs && hd && dt) SL(false, true, true);
    else SL(true, true, true);
#undef SL
    if (dt && out) {
        const int nvec = NV;
        hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(1024), 0, st, grid, nvec, partials, out);
    }
    return launch_status();
}


template <int NV>
int stream_for_nv(const StreamArgs& a) {
#define ST(TV, TX)                                                                              \
    do {                                                                                        \
        if (a.s32)                                                                              \
            return stream_launch<TV, TX, float, NV>(a.nb, a.blocks, a.slice_len, a.slice_base, \
                                                     a.rows,                                     \
                                                     a.col16, a.n,                               \
                                                     a.lg,                                       \
                                                     a.off, a.col,                               \
                                                     a.val, a.s, a.d, a.u, a.x, a.y, a.z,        \
                                                     a.alpha, a.beta, a.w, a.partials, a.out,    \
                                                     a.st);                                      \
        return stream_launch<TV, TX, TX, NV>(a.nb, a.blocks, a.slice_len, a.slice_base, a.rows,\
                                             a.col16,                                            \
                                             a.n, a.lg,                                          \
                                             a.off,                                              \
                                             a.col, a.val,                                       \
                                             a.s, a.d, a.u, a.x, a.y, a.z, a.alpha, a.beta, a.w, \
                                             a.partials, a.out, a.st);                           \
    } while (0)
    if (a.vec_dtype == WPC_F64) {
        if (a.val_dtype == WPC_F64) ST(double, double);
        if (a.val_dtype == WPC_F32) ST(float, double);
        if (a.val_dtype == WPC_F16) ST(f16s, double);
        if (a.val_dtype == WPC_BF16) ST(bf16s, double);
    } else if (a.vec_dtype == WPC_F32) {
        if (a.val_dtype == WPC_F32) ST(float, float);
        if (a.val_dtype == WPC_F16) ST(f16s, float);
        if (a.val_dtype == WPC_BF16) ST(bf16s, float);
    }
#undef ST
    return WPC_EDTYPE;
}

template <class TV, class TX, int NV>
int sqlap_cg_launch(const SqlapCgArgs& a) {
    const int grid = (int)blocks_for(a.n, kSqlapCgRows);  // each pass's partial count
#define SC(HW, FS, CF)                                                                          \
    do {                                                                                        \
        if ((a.passes & 1) && (a.passes & 8))                                                   \
        hipLaunchKernelGGL((sell_direction_energy_kernel<TV, TX, NV, HW, FS, CF>), dim3(grid),  \
                           dim3(kSqlapCgRows), 0, a.st, a.n, a.slice_ptr, a.slice_len,           \
                           a.slice_base, a.rows, a.col, (int)a.col16, (const TV*)a.val,         \
                           (const TX*)a.s, (const TX*)a.wm, (const TX*)a.z, (TX*)a.p, (TX*)a.t, \
                           a.P_rz, a.part_rz, grid, a.part_rr, a.state, a.part_e);              \
        else if (a.passes & 1)                                                                  \
        hipLaunchKernelGGL((sell_energy_kernel<TV, TX, NV, HW, CF>), dim3(grid),                \
                           dim3(kSqlapCgRows),                                                  \
                           0, a.st, a.n, a.slice_ptr, a.slice_len, a.slice_base, a.rows, a.col, \
                           (int)a.col16, (const TV*)a.val, (const TX*)a.s, (const TX*)a.wm,     \
                           (const TX*)a.p, (TX*)a.t, a.part_e);                                 \
        if (a.passes & 2)                                                                       \
        hipLaunchKernelGGL((sell_cg_update_kernel<TV, TX, NV, HW, FS, CF>), dim3(grid),         \
                           dim3(kSqlapCgRows), 0, a.st, a.n, a.slice_ptr, a.slice_len,           \
                           a.slice_base, a.rows, a.col, (int)a.col16, (const TV*)a.val,         \
                           (const TX*)a.wm, (const TX*)a.p, (const TX*)a.t, (TX*)a.r, grid,     \
                           a.part_e, a.state, a.part_rr, (TX*)a.x);                             \
    } while (0)
    const bool fs = (a.passes & 4) != 0;
    if (a.col16 == kColStencil) {
        if (a.wm != nullptr && fs) SC(true, true, kCfStencil);
        else if (a.wm != nullptr) SC(true, false, kCfStencil);
        else if (fs) SC(false, true, kCfStencil);
        else SC(false, false, kCfStencil);
    } else {
        if (a.wm != nullptr && fs) SC(true, true, kCfPlain);
        else if (a.wm != nullptr) SC(true, false, kCfPlain);
        else if (fs) SC(false, true, kCfPlain);
        else SC(false, false, kCfPlain);
    }
#undef SC
    return launch_status();
}

template <int NV>
int sqlap_cg_for_nv(const SqlapCgArgs& a) {
    if (a.n <= 0) return WPC_OK;
    if (a.vec_dtype == WPC_F64 && a.val_dtype == WPC_F64) return sqlap_cg_launch<double, double, NV>(a);
    if (a.vec_dtype == WPC_F32 && a.val_dtype == WPC_F32) return sqlap_cg_launch<float, float, NV>(a);
    return WPC_EDTYPE;
}

}  // namespace

#define WPC_STREAM_CAT2(a, b) a##b
#define WPC_STREAM_CAT(a, b) WPC_STREAM_CAT2(a, b)
#if WPC_STREAM_NV > 0
int WPC_STREAM_CAT(stream_entry_, WPC_STREAM_NV)(const StreamArgs& a) {
    return stream_for_nv<WPC_STREAM_NV>(a);
}
int WPC_STREAM_CAT(sqlap_cg_entry_, WPC_STREAM_NV)(const SqlapCgArgs& a) {
    return sqlap_cg_for_nv<WPC_STREAM_NV>(a);
}
#elif defined(WPC_GROUPED_F32)
int stream_grouped_f32(const GroupedArgs& a) { return grouped_for_nv<0, float>(a); }
#else
int stream_grouped_f32(const GroupedArgs& a);
int stream_grouped_entry(const GroupedArgs& a) {
    if (a.vec_dtype == WPC_F64) return grouped_for_nv<0, double>(a);
    if (a.vec_dtype == WPC_F32) return stream_grouped_f32(a);
    return WPC_EDTYPE;
}
#endif

}  // namespace wpc

#endif  // WPC_STREAM_NV
