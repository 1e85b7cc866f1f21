// Device assembly for the operators the preconditioners consume: P1 element
// matrices on tetrahedra, a deterministic segmented sum for triplet assembly, and
// the sparse product Q = K diag(1/m) K + alpha diag(m).
//
// The product follows the host pipeline it replaces term for term: for row i the
// contributions (K[i,k] * (1/m_k)) * K[k,j] are accumulated per output column over
// the stored k of row i in DESCENDING column order, then alpha*m_i is added to the
// diagonal.  That is the order of scipy's ((K @ diag(1/m)) @ K + alpha*diag(m)):
// its intermediate K @ diag(1/m) comes out of the row-wise linked-list product with
// each row's columns reversed, and the second product accumulates along that
// stored order.  One wavefront owns one output row; the row's columns live in an
// LDS open-addressing table.

#include "wpc_common.h"

namespace wpc {

// Per-tet gradients of the four P1 basis functions and |volume|; local stiffness
// vol * g_a . g_b (row-major 4x4) and the lumped mass share vol / 4.
__global__ __launch_bounds__(256) void p1_local_kernel(int ntet, const double* __restrict__ v,
                                                       const int* __restrict__ tets,
                                                       double* __restrict__ local,
                                                       double* __restrict__ quarter) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntet) return;
    int id[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) id[a] = tets[4L * t + a];
    double e[3][3];  // columns are edges v_a - v_0
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) e[c][a] = v[3L * id[a + 1] + c] - v[3L * id[0] + c];
    const double c00 = e[1][1] * e[2][2] - e[1][2] * e[2][1];
    const double c01 = e[1][2] * e[2][0] - e[1][0] * e[2][2];
    const double c02 = e[1][0] * e[2][1] - e[1][1] * e[2][0];
    const double det = e[0][0] * c00 + e[0][1] * c01 + e[0][2] * c02;
    // rows of the inverse are the gradients of phi_1..3
    double g[4][3];
    const double inv = 1.0 / det;
    g[1][0] = c00 * inv;
    g[1][1] = (e[0][2] * e[2][1] - e[0][1] * e[2][2]) * inv;
    g[1][2] = (e[0][1] * e[1][2] - e[0][2] * e[1][1]) * inv;
    g[2][0] = c01 * inv;
    g[2][1] = (e[0][0] * e[2][2] - e[0][2] * e[2][0]) * inv;
    g[2][2] = (e[0][2] * e[1][0] - e[0][0] * e[1][2]) * inv;
    g[3][0] = c02 * inv;
    g[3][1] = (e[0][1] * e[2][0] - e[0][0] * e[2][1]) * inv;
    g[3][2] = (e[0][0] * e[1][1] - e[0][1] * e[1][0]) * inv;
#pragma unroll
    for (int c = 0; c < 3; ++c) g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
    const double vol = fabs(det) / 6.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            local[16L * t + 4 * a + b] =
                vol * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
    quarter[t] = vol / 4.0;
}

// out[s] = sum of values[order[k]] for k in [seg[s], seg[s+1]), in k order.
__global__ __launch_bounds__(256) void segment_sum_kernel(int nseg, const long* __restrict__ seg,
                                                          const long* __restrict__ order,
                                                          const double* __restrict__ values,
                                                          double* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    double acc = 0.0;
    for (long k = seg[s]; k < seg[s + 1]; ++k) acc += values[order ? order[k] : k];
    out[s] = acc;
}

constexpr int kHash = 512;  // LDS slots per row (>= 2x the widest product row)
constexpr int kGemmWaves = 4;

__device__ __forceinline__ int hash_slot(int key) { return (int)(((unsigned)key * 2654435761u) % kHash); }

// Insert key (if absent) into the wave's table; returns its slot, or -1 when the
// table is full (bounded probing: every call terminates).
__device__ __forceinline__ int table_insert(int* keys, int key) {
    int h = hash_slot(key);
    for (int probe = 0; probe < kHash; ++probe) {
        const int prev = atomicCAS(keys + h, -1, key);
        if (prev == -1 || prev == key) return h;
        h = (h + 1) % kHash;
    }
    return -1;
}

// Symbolic pass: count[i] = number of distinct columns of row i of K*K plus the
// diagonal, or -1 when the row does not fit the table.
__global__ __launch_bounds__(256) void spgemm_count_kernel(int n, const int* __restrict__ off,
                                                           const int* __restrict__ col,
                                                           int* __restrict__ count) {
    __shared__ int keys_all[kGemmWaves][kHash];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int i = blockIdx.x * kGemmWaves + wave;
    if (i >= n) return;
    int* keys = keys_all[wave];
    for (int s = lane; s < kHash; s += kWave) keys[s] = -1;
    wave_sync();
    bool overflow = false;
    for (int e = off[i]; e < off[i + 1]; ++e) {
        const int k = col[e];
        for (int f = off[k] + lane; f < off[k + 1]; f += kWave)
            if (table_insert(keys, col[f]) < 0) overflow = true;
        wave_sync();
    }
    if (lane == 0 && table_insert(keys, i) < 0) overflow = true;
    wave_sync();
    int used = 0;
    for (int s = lane; s < kHash; s += kWave) used += keys[s] >= 0;
    for (int m = kWave / 2; m > 0; m >>= 1) used += __shfl_xor(used, m, kWave);
    overflow = __ballot(overflow) != 0;
    if (lane == 0) count[i] = overflow ? -1 : used;
}

// Numeric pass: fills columns (sorted) and values of row i at q_off[i].
__global__ __launch_bounds__(256) void spgemm_fill_kernel(
    int n, const int* __restrict__ off, const int* __restrict__ col, const double* __restrict__ val,
    const double* __restrict__ inv_m, const double* __restrict__ m, double alpha,
    const int* __restrict__ q_off, int* __restrict__ q_col, double* __restrict__ q_val) {
    __shared__ int keys_all[kGemmWaves][kHash];
    __shared__ double sums_all[kGemmWaves][kHash];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int i = blockIdx.x * kGemmWaves + wave;
    if (i >= n) return;
    int* keys = keys_all[wave];
    double* sums = sums_all[wave];
    for (int s = lane; s < kHash; s += kWave) {
        keys[s] = -1;
        sums[s] = 0.0;
    }
    wave_sync();
    for (int e = off[i + 1] - 1; e >= off[i]; --e) {
        const int k = col[e];
        const double a = val[e] * inv_m[k];  // entry of K @ diag(1/m)
        // distinct j within row k: each lane owns different slots in this step
        for (int f = off[k] + lane; f < off[k + 1]; f += kWave) {
            const int h = table_insert(keys, col[f]);  // fits: the count pass checked
            if (h >= 0) sums[h] += a * val[f];
        }
        wave_sync();
    }
    if (lane == 0) {
        const int h = table_insert(keys, i);
        if (h >= 0) sums[h] += alpha * m[i];
    }
    wave_sync();
    // rank-sort the occupied slots by column
    const int base = q_off[i];
    for (int s = lane; s < kHash; s += kWave) {
        const int key = keys[s];
        if (key < 0) continue;
        int rank = 0;
        for (int t = 0; t < kHash; ++t) rank += (keys[t] >= 0 && keys[t] < key);
        q_col[base + rank] = key;
        q_val[base + rank] = sums[s];
    }
}

}  // namespace wpc

using namespace wpc;

extern "C" {

int wpc_p1_local(int ntet, const double* vertices, const int* tets, double* local,
                 double* quarter_volume, void* stream) {
    if (ntet <= 0) return WPC_OK;
    hipLaunchKernelGGL(p1_local_kernel, dim3(blocks_for(ntet, 256)), dim3(256), 0,
                       as_stream(stream), ntet, vertices, tets, local, quarter_volume);
    return launch_status();
}

int wpc_segment_sum(int nseg, const long* segments, const long* order, const double* values,
                    double* out, void* stream) {
    if (nseg <= 0) return WPC_OK;
    hipLaunchKernelGGL(segment_sum_kernel, dim3(blocks_for(nseg, 256)), dim3(256), 0,
                       as_stream(stream), nseg, segments, order, values, out);
    return launch_status();
}

int wpc_spgemm_kmk_count(int n, const int* offsets, const int* columns, int* counts,
                         void* stream) {
    if (n <= 0) return WPC_OK;
    hipLaunchKernelGGL(spgemm_count_kernel, dim3(blocks_for(n, kGemmWaves)),
                       dim3(kGemmWaves * kWave), 0, as_stream(stream), n, offsets, columns, counts);
    return launch_status();
}

int wpc_spgemm_kmk_fill(int n, const int* offsets, const int* columns, const double* values,
                        const double* inverse_mass, const double* mass, double alpha,
                        const int* q_offsets, int* q_columns, double* q_values, void* stream) {
    if (n <= 0) return WPC_OK;
    hipLaunchKernelGGL(spgemm_fill_kernel, dim3(blocks_for(n, kGemmWaves)),
                       dim3(kGemmWaves * kWave), 0, as_stream(stream), n, offsets, columns, values,
                       inverse_mass, mass, alpha, q_offsets, q_columns, q_values);
    return launch_status();
}

}  // extern "C"
