// Device routines shared by the FSAI build and refit kernels.
#pragma once

#include "wpc_common.h"

namespace wpc {

// fsai.py:20-41 (entry): A[i, j] by binary search of a sorted row, zero if absent.
template <class T, class S>
__device__ __forceinline__ T entry(const int* __restrict__ off, const int* __restrict__ col,
                                   const S* __restrict__ val, int i, int j) {
    int lo = off[i], hi = off[i + 1];
    const int end = hi;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    T result = T(0);
    if (lo < end && col[lo] == j) result = ldv<T, S>(val, lo);
    return result;
}

// entry() for up to Q targets j[0..cnt) in one row i at once: the searches advance
// in lockstep (each level's loads of all targets in one round trip instead of one
// dependent chain per target) and each returns exactly what entry() returns.
template <class T, class S, int Q>
__device__ __forceinline__ void entries(const int* __restrict__ off, const int* __restrict__ col,
                                        const S* __restrict__ val, int i, const int* j, int cnt,
                                        T* out) {
    const int start = off[i], end = off[i + 1];
    int lo[Q], hi[Q], tgt[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        tgt[q] = q < cnt ? j[q] : 0;
        lo[q] = start;
        hi[q] = q < cnt ? end : start;
    }
    bool busy = start < end && cnt > 0;
    while (busy) {
        busy = false;
        int c[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int mid = (lo[q] + hi[q]) >> 1;
            c[q] = lo[q] < hi[q] ? col[mid] : 0;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (lo[q] < hi[q]) {
                const int mid = (lo[q] + hi[q]) >> 1;
                if (c[q] < tgt[q]) lo[q] = mid + 1;
                else hi[q] = mid;
                busy = busy || lo[q] < hi[q];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        T result = T(0);
        if (q < cnt && lo[q] < end && col[lo[q]] == tgt[q]) result = ldv<T, S>(val, lo[q]);
        out[q] = result;
    }
}

// Packed lower-triangular index (row r >= column c).
__host__ __device__ inline int tri(int r, int c) { return r * (r + 1) / 2 + c; }

}  // namespace wpc
