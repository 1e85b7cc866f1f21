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

// Packed lower-triangular index (row r >= column c).
__host__ __device__ inline int tri(int r, int c) { return r * (r + 1) / 2 + c; }

}  // namespace wpc
