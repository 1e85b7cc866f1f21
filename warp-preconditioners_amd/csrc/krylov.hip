// Vector steps of batched preconditioned CG and CR with device-resident scalars.
//
// nvec right-hand sides advance together on interleaved [n][nvec] vectors.  Every
// inner product is reduced deterministically: each of P blocks writes one partial
// per vector (fixed tree order), and the consuming kernel folds the P partials in
// a fixed order in every block, so no host round trip and no atomics are needed.
// A slot whose incoming residual already meets its target takes no step, which
// lets a captured chunk of iterations run past convergence unchanged.

#include "reduce.h"

namespace wpc {

constexpr int kKrylovThreads = kStreamRows;

int krylov_blocks(int n) {
    long b = ((long)n + 511) / 512;
    if (b > 1024) b = 1024;
    return (int)(b < 1 ? 1 : b);
}

// Fold partials[0..P) (stride nvec) for vector v with all threads of the block, in
// a fixed order; the result lands in red[0].
template <int NV>
__device__ void fold(const double* __restrict__ partials, int P, int v, double* red) {
    double s = 0.0;
    for (int b = threadIdx.x; b < P; b += kKrylovThreads) s += partials[(long)b * NV + v];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kKrylovThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
}

// Block-level sum of per-thread values acc[NV] -> partials[blockIdx.x*NV + v].
template <int NV>
__device__ void block_partial(const double* acc, double* red, double* partials) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        red[threadIdx.x] = acc[v];
        __syncthreads();
        for (int w = kKrylovThreads / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(long)blockIdx.x * NV + v] = red[0];
        __syncthreads();
    }
}

__device__ __forceinline__ void chunk_of(int n, int blk, int nblocks, long* begin, long* end) {
    // (32-bit unsigned: n < 2^31, so n + nblocks - 1 < 2^32; a 64-bit division
    // here costs a long scalar sequence ahead of every load)
    const unsigned chunk = ((unsigned)n + (unsigned)nblocks - 1u) / (unsigned)nblocks;
    *begin = (long)blk * chunk;
    *end = *begin + chunk < n ? *begin + chunk : n;
}

__device__ __forceinline__ void chunk_of(int n, long* begin, long* end) {
    chunk_of(n, blockIdx.x, gridDim.x, begin, end);
}

template <class T, int NV>
__global__ __launch_bounds__(kKrylovThreads) void kdot_kernel(int n, const T* __restrict__ a,
                                                              const T* __restrict__ b,
                                                              double* __restrict__ partials) {
    __shared__ double red[kKrylovThreads];
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    long begin, end;
    chunk_of(n, &begin, &end);
    for (long i = begin + threadIdx.x; i < end; i += kKrylovThreads) {
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += (double)a[i * NV + v] * (double)b[i * NV + v];
    }
    block_partial<NV>(acc, red, partials);
}

// alpha = active ? rz / denom : 0;  x += alpha p;  r -= alpha Ap;  (CR: z -= alpha MAp)
template <class T, int NV, bool CR>
__global__ __launch_bounds__(kKrylovThreads) void kupdate_kernel(
    int n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ ap,
    T* __restrict__ z, const T* __restrict__ map, double* __restrict__ state,
    const double* __restrict__ part_denom, double* __restrict__ part_rr) {
    __shared__ double red[kKrylovThreads];
    __shared__ double alpha_s[NV];
    const int P = gridDim.x;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        fold<NV>(part_denom, P, v, red);
        if (threadIdx.x == 0) {
            const double denom = red[0];
            const double rz = state[WPC_KS_RZ * NV + v];
            const bool active = state[WPC_KS_RR * NV + v] > state[WPC_KS_TARGET * NV + v];
            const double alpha = active ? rz / denom : 0.0;
            alpha_s[v] = alpha;
            if (blockIdx.x == 0) {
                state[WPC_KS_ALPHA * NV + v] = alpha;
                state[WPC_KS_DENOM * NV + v] = denom;
                state[WPC_KS_ACTIVE * NV + v] = active ? 1.0 : 0.0;
                state[WPC_KS_RZ_OLD * NV + v] = rz;
            }
        }
        __syncthreads();
    }
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    long begin, end;
    chunk_of(n, &begin, &end);
    for (long i = begin + threadIdx.x; i < end; i += kKrylovThreads) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long k = i * NV + v;
            const T a = (T)alpha_s[v];
            T rk = r[k];
            if (a != T(0)) {
                x[k] += a * p[k];
                rk -= a * ap[k];
                r[k] = rk;
                if (CR) z[k] -= a * map[k];
            }
            acc[v] += (double)rk * (double)rk;
        }
    }
    block_partial<NV>(acc, red, part_rr);
}

// beta = rz_new / rz_old; p = z + beta p (CR: Ap = Az + beta Ap) for active slots.
template <class T, int NV, bool CR>
__global__ __launch_bounds__(kKrylovThreads) void kdirection_kernel(
    int n, const T* __restrict__ z, T* __restrict__ p, const T* __restrict__ az,
    T* __restrict__ ap, double* __restrict__ state, const double* __restrict__ part_rz,
    const double* __restrict__ part_rr) {
    __shared__ double red[kKrylovThreads];
    __shared__ double beta_s[NV];
    __shared__ int active_s[NV];
    const int P = gridDim.x;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        fold<NV>(part_rz, P, v, red);
        double rz_new = red[0];
        __syncthreads();
        fold<NV>(part_rr, P, v, red);
        if (threadIdx.x == 0) {
            const double rr = red[0];
            const bool active = state[WPC_KS_ACTIVE * NV + v] != 0.0;
            const double rz_old = state[WPC_KS_RZ_OLD * NV + v];
            beta_s[v] = active ? rz_new / rz_old : 0.0;
            active_s[v] = active;
            if (blockIdx.x == 0 && active) {
                state[WPC_KS_RZ * NV + v] = rz_new;
                state[WPC_KS_RR * NV + v] = rr;
                state[WPC_KS_ITER * NV + v] += 1.0;
            }
        }
        __syncthreads();
    }
    long begin, end;
    chunk_of(n, &begin, &end);
    for (long i = begin + threadIdx.x; i < end; i += kKrylovThreads) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (!active_s[v]) continue;
            const long k = i * NV + v;
            const T b = (T)beta_s[v];
            p[k] = z[k] + b * p[k];
            if (CR) ap[k] = az[k] + b * ap[k];
        }
    }
}


// ---------------------------------------------------------------- fused CG steps
// The operator's last pass leaves p.Ap partials and the preconditioner's last
// pass r.z partials (wpc_csr_stream with a dot and out = NULL); two
// single-workgroup scalar launches fold them, so a CG iteration is
//   A p (+partials) | scalars_alpha | update (+r.r partials) | M r (+partials) |
//   scalars_beta | direction.

// DENOM = fold(p.Ap); ACTIVE = RR > TARGET; ALPHA = ACTIVE ? RZ / DENOM : 0.
__global__ __launch_bounds__(1024) void cg_alpha_kernel(
    int nvec, int P, const double* __restrict__ part_denom, double* __restrict__ state) {
    __shared__ double sums[8];
    fold_all<1024>(part_denom, P, nvec, sums);
    const int v = threadIdx.x;
    {
        if (v < nvec) {
            const double denom = sums[v];
            const bool active = state[WPC_KS_RR * nvec + v] > state[WPC_KS_TARGET * nvec + v];
            state[WPC_KS_DENOM * nvec + v] = denom;
            state[WPC_KS_ACTIVE * nvec + v] = active ? 1.0 : 0.0;
            state[WPC_KS_ALPHA * nvec + v] = active ? state[WPC_KS_RZ * nvec + v] / denom : 0.0;
        }
    }
}

// x += ALPHA p;  r -= ALPHA Ap;  partials of r.r.
template <class T, int NV>
__global__ __launch_bounds__(kKrylovThreads) void cg_update_kernel(
    int n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
    const T* __restrict__ ap, const double* __restrict__ state, double* __restrict__ partials) {
    __shared__ double red[kKrylovThreads];
    T alpha[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) alpha[v] = (T)state[WPC_KS_ALPHA * NV + v];
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    long begin, end;
    chunk_of(n, &begin, &end);
    for (long i = begin + threadIdx.x; i < end; i += kKrylovThreads) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long k = i * NV + v;
            T rk = r[k];
            if (alpha[v] != T(0)) {
                x[k] += alpha[v] * p[k];
                rk -= alpha[v] * ap[k];
                r[k] = rk;
            }
            acc[v] += (double)rk * (double)rk;
        }
    }
    block_sum<NV>(acc, red);
    write_partials<NV>(acc, partials);
}

// RZ_NEW = fold(r.z), RR_NEW = fold(r.r); for stepping slots BETA = RZ_NEW / RZ,
// then RZ <- RZ_NEW, RR <- RR_NEW, ITER += 1 (BETA = 0 and nothing committed
// for slots that did not step).
__global__ __launch_bounds__(1024) void cg_beta_kernel(
    int nvec, int P_rz, const double* __restrict__ part_rz, int P_rr,
    const double* __restrict__ part_rr, double* __restrict__ state) {
    __shared__ double s_rz[8], s_rr[8];
    fold_two<1024>(part_rz, P_rz, part_rr, P_rr, nvec, s_rz, s_rr);
    const int v = threadIdx.x;
    {
        if (v < nvec) {
            const double rz_new = s_rz[v];
            const double rr = s_rr[v];
            const bool active = state[WPC_KS_ACTIVE * nvec + v] != 0.0;
            state[WPC_KS_RZ_NEW * nvec + v] = rz_new;
            state[WPC_KS_RR_NEW * nvec + v] = rr;
            if (active) {
                state[WPC_KS_BETA * nvec + v] = rz_new / state[WPC_KS_RZ * nvec + v];
                state[WPC_KS_RZ * nvec + v] = rz_new;
                state[WPC_KS_RR * nvec + v] = rr;
                state[WPC_KS_ITER * nvec + v] += 1.0;
            } else {
                state[WPC_KS_BETA * nvec + v] = 0.0;
            }
        }
    }
}

// The same two steps with the scalar folds folded in, for partial counts small
// enough that every workgroup can afford to fold them: each workgroup folds in the
// same fixed order, so all agree bit for bit, and workgroup 0 records the scalars.
// The fields each kernel reads and the fields workgroup 0 writes are disjoint:
//   update:    reads RZ, RR, TARGET      writes RZ_OLD (= RZ), DENOM, ACTIVE, ALPHA
//   direction: reads RZ_OLD, ACTIVE      writes RZ_NEW, RR_NEW, BETA, RZ, RR, ITER
// Rows each thread loads before the fold, so that the HBM reads overlap the fold's
// latency (bounded so the prefetched values stay in registers).
template <int NV> constexpr int kPrefetch = NV <= 2 ? 4 : (NV <= 4 ? 2 : 1);

// XD: x is advanced by the direction step (x += ALPHA p there, before p is
// replaced), so this step reads only r and Ap and leaves x and p alone: one
// vector stream fewer per iteration than advancing x here, where p must be read
// for x alone.  Each x element takes the same operation either way.
template <class T, int NV, bool XD>
__device__ __forceinline__ void cg_update_fold_body(
    int blk, int nblocks, int n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
    const T* __restrict__ ap, int P, const double* __restrict__ part_denom, double* state,
    double* __restrict__ partials, double* red, double* sums, int nv_rt = NV) {
    constexpr int PF = NV > 0 ? kPrefetch<NV> : 1;
    constexpr int NVM = NV > 0 ? NV : 8;
    const int nv = NV > 0 ? NV : nv_rt;
    long begin, end;
    chunk_of(n, blk, nblocks, &begin, &end);
    // Loads in flight together: the state fields read, the partials (compile-time
    // counts), then the prefetched rows.  The fields read and written are disjoint,
    // so reading them before workgroup 0's writes changes nothing.
    StateLanes<WPC_KS_RZ, WPC_KS_RR, WPC_KS_TARGET> st;
    st.load(state, nv);
    FoldRegs<kKrylovThreads, (NV > 0 ? NV : 1)> fold_a;
    if constexpr (NV > 0) fold_a.load(part_denom, P);
    T xv[PF][NVM], rv[PF][NVM], pv[PF][NVM], av[PF][NVM];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        const long i = begin + threadIdx.x + (long)j * kKrylovThreads;
        if (i < end) {
#pragma unroll
            for (int v = 0; v < NVM; ++v) if (v < nv) {
                const long k = i * nv + v;
                if constexpr (!XD) {
                    xv[j][v] = x[k];
                    pv[j][v] = p[k];
                }
                rv[j][v] = r[k];
                av[j][v] = ap[k];
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);  // (as in the direction step below)
    if constexpr (NV > 0)
        fold_a.finish(part_denom, P, sums);
    else
        fold_all<kKrylovThreads>(part_denom, P, nv, sums);
    T alpha[NVM];
#pragma unroll
    for (int v = 0; v < NVM; ++v) if (v < nv) {
        const double denom = sums[v];
        const double rz = st.get(0, nv, v);
        const bool active = st.get(1, nv, v) > st.get(2, nv, v);
        const double a = active ? rz / denom : 0.0;
        alpha[v] = (T)a;
        if (blk == 0 && threadIdx.x == v) {
            state[WPC_KS_RZ_OLD * nv + v] = rz;
            state[WPC_KS_DENOM * nv + v] = denom;
            state[WPC_KS_ACTIVE * nv + v] = active ? 1.0 : 0.0;
            state[WPC_KS_ALPHA * nv + v] = a;
        }
    }
    double acc[NVM];
    bool moves = false;  // some slot steps: rows are stored whole (a resting slot's
#pragma unroll           // values unchanged), so the stores of a row merge
    for (int v = 0; v < NVM; ++v) if (v < nv) {
        acc[v] = 0.0;
        moves = moves || alpha[v] != T(0);
    }
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        const long i = begin + threadIdx.x + (long)j * kKrylovThreads;
        if (i < end) {
            T xn[NVM], rn[NVM];
#pragma unroll
            for (int v = 0; v < NVM; ++v) if (v < nv) {
                const bool step = alpha[v] != T(0);
                if constexpr (!XD) xn[v] = step ? xv[j][v] + alpha[v] * pv[j][v] : xv[j][v];
                rn[v] = step ? rv[j][v] - alpha[v] * av[j][v] : rv[j][v];
                acc[v] += (double)rn[v] * (double)rn[v];
            }
            if (moves) {
#pragma unroll
                for (int v = 0; v < NVM; ++v) if (v < nv) {
                    if constexpr (!XD) x[i * nv + v] = xn[v];
                    r[i * nv + v] = rn[v];
                }
            }
        }
    }
    for (long i = begin + threadIdx.x + (long)PF * kKrylovThreads; i < end; i += kKrylovThreads) {
#pragma unroll
        for (int v = 0; v < NVM; ++v) if (v < nv) {
            const long k = i * nv + v;
            T rk = r[k];
            if (alpha[v] != T(0)) {
                if constexpr (!XD) x[k] += alpha[v] * p[k];
                rk -= alpha[v] * ap[k];
                r[k] = rk;
            }
            acc[v] += (double)rk * (double)rk;
        }
    }
    if constexpr (NV > 0) {
        block_sum<NV>(acc, red);
        write_partials<NV>(acc, partials, blk);
    } else {
        block_sum_n(acc, nv, red);
        write_partials_n(acc, nv, partials, blk);
    }
}

template <class T, int NV, bool XD>
__global__ __launch_bounds__(kKrylovThreads) void cg_update_fold_kernel(
    int n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
    const T* __restrict__ ap, int P, const double* __restrict__ part_denom, double* state,
    double* __restrict__ partials) {
    __shared__ double red[kKrylovThreads];
    __shared__ double sums[8];
    cg_update_fold_body<T, NV, XD>(blockIdx.x, gridDim.x, n, x, r, p, ap, P, part_denom, state,
                                   partials, red, sums);
}

// XD: also x += ALPHA p (the update's step, for an update that left x alone),
// from the p this step replaces.
template <class T, int NV, bool XD>
__device__ __forceinline__ void cg_direction_fold_body(
    int blk, int nblocks, int n, const T* __restrict__ z, T* __restrict__ p, T* __restrict__ x,
    int P_rz, const double* __restrict__ part_rz, int P_rr, const double* __restrict__ part_rr,
    double* state, double* s_rz, double* s_rr, int nv_rt = NV) {
    constexpr int PF = NV > 0 ? kPrefetch<NV> : 1;
    constexpr int NVM = NV > 0 ? NV : 8;
    const int nv = NV > 0 ? NV : nv_rt;
    long begin, end;
    chunk_of(n, blk, nblocks, &begin, &end);
    // As in the update: state fields, partials, then rows, all in flight together.
    StateLanes<WPC_KS_ACTIVE, WPC_KS_RZ_OLD, (XD ? WPC_KS_ALPHA : -1)> st;
    st.load(state, nv);
    FoldRegs<kKrylovThreads, (NV > 0 ? NV : 1)> fold_z, fold_r;
    if constexpr (NV > 0) {
        fold_z.load(part_rz, P_rz);
        if (blk == 0) fold_r.load(part_rr, P_rr);  // r.r is only recorded, by workgroup 0
    }
    T zv[PF][NVM], pv[PF][NVM], xv[PF][NVM];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        const long i = begin + threadIdx.x + (long)j * kKrylovThreads;
        if (i < end) {
#pragma unroll
            for (int v = 0; v < NVM; ++v) if (v < nv) {
                zv[j][v] = z[i * nv + v];
                pv[j][v] = p[i * nv + v];
                if constexpr (XD) xv[j][v] = x[i * nv + v];
            }
        }
    }
    // every load above goes out before the first wait: the scheduler would start
    // the folds' adds early (to free registers) and wait for each row in turn
    __builtin_amdgcn_sched_barrier(0);
    T alpha[NVM];
    if constexpr (XD) {
        // x += ALPHA p needs no fold: its stores go out while the folds run
        bool steps = false;
#pragma unroll
        for (int v = 0; v < NVM; ++v) if (v < nv) {
            alpha[v] = (T)st.get(2, nv, v);
            steps = steps || alpha[v] != T(0);
        }
        if (steps) {
#pragma unroll
            for (int j = 0; j < PF; ++j) {
                const long i = begin + threadIdx.x + (long)j * kKrylovThreads;
                if (i < end) {
#pragma unroll
                    for (int v = 0; v < NVM; ++v) if (v < nv)
                        x[i * nv + v] = alpha[v] != T(0) ? xv[j][v] + alpha[v] * pv[j][v]
                                                         : xv[j][v];
                }
            }
        }
    }
    if constexpr (NV > 0) {
        if (blk != 0) {
            fold_z.finish(part_rz, P_rz, s_rz);
        } else if constexpr (NV <= 4) {
            // both folds through one cross-wavefront pass (one barrier pair, not two)
            double acc[8];
            __shared__ double both[8];
            fold_z.sums(part_rz, P_rz, acc);
            fold_r.sums(part_rr, P_rr, acc + NV);
            fold_reduce<kKrylovThreads>(acc, 2 * NV, both);
#pragma unroll
            for (int v = 0; v < NV; ++v) { s_rz[v] = both[v]; s_rr[v] = both[NV + v]; }
        } else {
            fold_z.finish(part_rz, P_rz, s_rz);
            fold_r.finish(part_rr, P_rr, s_rr);
        }
    } else {
        fold_all<kKrylovThreads>(part_rz, P_rz, nv, s_rz);
        if (blk == 0) fold_all<kKrylovThreads>(part_rr, P_rr, nv, s_rr);
    }
    T beta[NVM];
    bool act[NVM];
#pragma unroll
    for (int v = 0; v < NVM; ++v) if (v < nv) {
        act[v] = st.get(0, nv, v) != 0.0;
        const double b = act[v] ? s_rz[v] / st.get(1, nv, v) : 0.0;
        beta[v] = (T)b;
        if (blk == 0 && threadIdx.x == v) {
            state[WPC_KS_RZ_NEW * nv + v] = s_rz[v];
            state[WPC_KS_RR_NEW * nv + v] = s_rr[v];
            state[WPC_KS_BETA * nv + v] = b;
            if (act[v]) {
                state[WPC_KS_RZ * nv + v] = s_rz[v];
                state[WPC_KS_RR * nv + v] = s_rr[v];
                state[WPC_KS_ITER * nv + v] += 1.0;
            }
        }
    }
    bool moves = false;  // as in the update: whole rows, stored when some slot steps
#pragma unroll
    for (int v = 0; v < NVM; ++v) if (v < nv) moves = moves || act[v];
    if (moves) {
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            const long i = begin + threadIdx.x + (long)j * kKrylovThreads;
            if (i < end) {
#pragma unroll
                for (int v = 0; v < NVM; ++v) if (v < nv)
                    p[i * nv + v] = act[v] ? zv[j][v] + beta[v] * pv[j][v] : pv[j][v];
            }
        }
    }
    for (long i = begin + threadIdx.x + (long)PF * kKrylovThreads; i < end; i += kKrylovThreads) {
#pragma unroll
        for (int v = 0; v < NVM; ++v) if (v < nv) {
            if (!act[v]) continue;
            const long k = i * nv + v;
            const T pk = p[k];
            if constexpr (XD)
                if (alpha[v] != T(0)) x[k] += alpha[v] * pk;
            p[k] = z[k] + beta[v] * pk;
        }
    }
}

template <class T, int NV, bool XD>
__global__ __launch_bounds__(kKrylovThreads) void cg_direction_fold_kernel(
    int n, const T* __restrict__ z, T* __restrict__ p, T* __restrict__ x, int P_rz,
    const double* __restrict__ part_rz, int P_rr, const double* __restrict__ part_rr,
    double* state) {
    __shared__ double s_rz[8], s_rr[8];
    cg_direction_fold_body<T, NV, XD>(blockIdx.x, gridDim.x, n, z, p, x, P_rz, part_rz, P_rr,
                                      part_rr, state, s_rz, s_rr);
}

// One CG vector step of a grouped launch (all fields 64-bit, in this order):
// vectors, the partials the step folds (a: p.Ap or r.z; b: r.r), the state, the
// r.r partials it leaves (update), the system's size, its workgroup count and
// first global workgroup, and the folded partial counts.
struct CgTask {
    void* x;
    void* r;
    void* p;
    const void* ap;
    const void* z;
    const double* part_a;
    const double* part_b;
    double* state;
    double* partials;
    long n;
    long nblocks;
    long block0;
    long pa;
    long pb;
    long nv;      // interleaved vectors of this system (1..8)
    long unused;  // tables of both kinds are 16 words wide
};
static_assert(sizeof(CgTask) == 16 * 8, "CgTask is 16 64-bit words");

__device__ __forceinline__ int task_of(const CgTask* __restrict__ tasks, int ntasks, int g,
                                       int* s_task) {
    if (threadIdx.x == 0) {
        int lo = 0, hi = ntasks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (tasks[mid].block0 <= g) lo = mid;
            else hi = mid - 1;
        }
        *s_task = lo;
    }
    __syncthreads();
    return *s_task;
}

// The fused update / direction steps of several systems in one launch: each
// workgroup runs its task's local block, folding that task's partials, exactly as
// the system's own launch would.
template <class T, int NV>
__global__ __launch_bounds__(kKrylovThreads) void grouped_update_fold_kernel(
    const CgTask* __restrict__ tasks, int ntasks) {
    __shared__ double red[kKrylovThreads];
    __shared__ double sums[8];
    __shared__ int s_task;
    const int g = blockIdx.x;
    const CgTask& t = tasks[task_of(tasks, ntasks, g, &s_task)];
    if (t.n == 0) return;  // capacity padding
    cg_update_fold_body<T, NV, true>(g - (int)t.block0, (int)t.nblocks, (int)t.n, (T*)t.x,
                                     (T*)t.r, (const T*)t.p, (const T*)t.ap, (int)t.pa,
                                     t.part_a, t.state, t.partials, red, sums, (int)t.nv);
}

template <class T, int NV>
__global__ __launch_bounds__(kKrylovThreads) void grouped_direction_fold_kernel(
    const CgTask* __restrict__ tasks, int ntasks) {
    __shared__ double s_rz[8], s_rr[8];
    __shared__ int s_task;
    const int g = blockIdx.x;
    const CgTask& t = tasks[task_of(tasks, ntasks, g, &s_task)];
    if (t.n == 0) return;  // capacity padding
    cg_direction_fold_body<T, NV, true>(g - (int)t.block0, (int)t.nblocks, (int)t.n,
                                        (const T*)t.z, (T*)t.p, (T*)t.x, (int)t.pa, t.part_a,
                                        (int)t.pb, t.part_b, t.state, s_rz, s_rr, (int)t.nv);
}

// p = z + BETA p for slots that stepped.
template <class T, int NV>
__global__ __launch_bounds__(kKrylovThreads) void cg_direction_kernel(
    int n, const T* __restrict__ z, T* __restrict__ p, const double* __restrict__ state) {
    T beta[NV];
    bool act[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        act[v] = state[WPC_KS_ACTIVE * NV + v] != 0.0;
        beta[v] = (T)state[WPC_KS_BETA * NV + v];
    }
    long begin, end;
    chunk_of(n, &begin, &end);
    for (long i = begin + threadIdx.x; i < end; i += kKrylovThreads) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (!act[v]) continue;
            const long k = i * NV + v;
            p[k] = z[k] + beta[v] * p[k];
        }
    }
}

// target = max(atol, tol * sqrt(||b||^2))^2 in place (state[TARGET] holds ||b||^2).
__global__ void ktarget_kernel(int nvec, double* state, double tol, double atol) {
    const int v = threadIdx.x;
    if (v >= nvec) return;
    const double bnorm = sqrt(state[WPC_KS_TARGET * nvec + v]);
    const double t = fmax(atol, tol * bnorm);
    state[WPC_KS_TARGET * nvec + v] = t * t;
    state[WPC_KS_ITER * nvec + v] = 0.0;
}

}  // namespace wpc

using namespace wpc;

#define NV_SWITCH(nvec, CALL)                               \
    switch (nvec) {                                         \
        case 1: { constexpr int NV = 1; CALL; break; }      \
        case 2: { constexpr int NV = 2; CALL; break; }      \
        case 3: { constexpr int NV = 3; CALL; break; }      \
        case 4: { constexpr int NV = 4; CALL; break; }      \
        case 5: { constexpr int NV = 5; CALL; break; }      \
        case 6: { constexpr int NV = 6; CALL; break; }      \
        case 7: { constexpr int NV = 7; CALL; break; }      \
        case 8: { constexpr int NV = 8; CALL; break; }      \
        default: return WPC_EARG;                           \
    }
#define VT_SWITCH(dt, CALL)                                 \
    if ((dt) == WPC_F64) { using T = double; CALL; }        \
    else if ((dt) == WPC_F32) { using T = float; CALL; }    \
    else return WPC_EDTYPE;

extern "C" {

int wpc_krylov_partials(int n) { return krylov_blocks(n); }

int wpc_krylov_dot(int n, int nvec, const void* a, const void* b, int dtype, double* partials,
                   void* stream) {
    const int P = krylov_blocks(n);
    hipStream_t s = as_stream(stream);
    VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((kdot_kernel<T, NV>), dim3(P), dim3(kKrylovThreads), 0, s, n, (const T*)a, (const T*)b, partials)))
    return launch_status();
}

int wpc_krylov_update(int method, int n, int nvec, void* x, void* r, const void* p,
                      const void* ap, void* z, const void* map, int dtype, double* state,
                      const double* part_denom, double* part_rr, void* stream) {
    const int P = krylov_blocks(n);
    hipStream_t s = as_stream(stream);
    if (method == 0) {
        VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((kupdate_kernel<T, NV, false>), dim3(P), dim3(kKrylovThreads), 0, s, n, (T*)x, (T*)r, (const T*)p, (const T*)ap, (T*)z, (const T*)map, state, part_denom, part_rr)))
    } else {
        VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((kupdate_kernel<T, NV, true>), dim3(P), dim3(kKrylovThreads), 0, s, n, (T*)x, (T*)r, (const T*)p, (const T*)ap, (T*)z, (const T*)map, state, part_denom, part_rr)))
    }
    return launch_status();
}

int wpc_krylov_direction(int method, int n, int nvec, const void* z, void* p, const void* az,
                         void* ap, int dtype, double* state, const double* part_rz,
                         const double* part_rr, void* stream) {
    const int P = krylov_blocks(n);
    hipStream_t s = as_stream(stream);
    if (method == 0) {
        VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((kdirection_kernel<T, NV, false>), dim3(P), dim3(kKrylovThreads), 0, s, n, (const T*)z, (T*)p, (const T*)az, (T*)ap, state, part_rz, part_rr)))
    } else {
        VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((kdirection_kernel<T, NV, true>), dim3(P), dim3(kKrylovThreads), 0, s, n, (const T*)z, (T*)p, (const T*)az, (T*)ap, state, part_rz, part_rr)))
    }
    return launch_status();
}

int wpc_krylov_set_target(int nvec, double* state, double tol, double atol, void* stream) {
    hipLaunchKernelGGL(ktarget_kernel, dim3(1), dim3(64), 0, as_stream(stream), nvec, state, tol, atol);
    return launch_status();
}

int wpc_cg_alpha(int nvec, int nparts, const double* part_denom, double* state, void* stream) {
    hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(1024), 0, as_stream(stream), nvec,
                       nparts, part_denom, state);
    return launch_status();
}

int wpc_cg_update(int n, int nvec, void* x, void* r, const void* p, const void* ap, int dtype,
                  const double* state, double* partials, void* stream) {
    const int P = krylov_blocks(n);
    hipStream_t s = as_stream(stream);
    VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((cg_update_kernel<T, NV>), dim3(P), dim3(kKrylovThreads), 0, s, n, (T*)x, (T*)r, (const T*)p, (const T*)ap, state, partials)))
    return launch_status();
}

int wpc_cg_beta(int nvec, int nparts_rz, const double* part_rz, int nparts_rr,
                const double* part_rr, double* state, void* stream) {
    hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(1024), 0, as_stream(stream), nvec,
                       nparts_rz, part_rz, nparts_rr, part_rr, state);
    return launch_status();
}

int wpc_cg_direction(int n, int nvec, const void* z, void* p, int dtype, const double* state,
                     void* stream) {
    const int P = krylov_blocks(n);
    hipStream_t s = as_stream(stream);
    VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((cg_direction_kernel<T, NV>), dim3(P), dim3(kKrylovThreads), 0, s, n, (const T*)z, (T*)p, state)))
    return launch_status();
}

int wpc_cg_update_fold(int n, int nvec, void* x, void* r, const void* p, const void* ap, int dtype,
                       int nparts, const double* part_denom, double* state, double* partials,
                       void* stream) {
    const int P = krylov_blocks(n);
    hipStream_t s = as_stream(stream);
    if (x == nullptr) {
        VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((cg_update_fold_kernel<T, NV, true>), dim3(P), dim3(kKrylovThreads), 0, s, n, (T*)x, (T*)r, (const T*)p, (const T*)ap, nparts, part_denom, state, partials)))
    } else {
        VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((cg_update_fold_kernel<T, NV, false>), dim3(P), dim3(kKrylovThreads), 0, s, n, (T*)x, (T*)r, (const T*)p, (const T*)ap, nparts, part_denom, state, partials)))
    }
    return launch_status();
}

int wpc_cg_direction_fold(int n, int nvec, const void* z, void* p, void* x, int dtype,
                          int nparts_rz, const double* part_rz, int nparts_rr,
                          const double* part_rr, double* state, void* stream) {
    const int P = krylov_blocks(n);
    hipStream_t s = as_stream(stream);
    if (x != nullptr) {
        VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((cg_direction_fold_kernel<T, NV, true>), dim3(P), dim3(kKrylovThreads), 0, s, n, (const T*)z, (T*)p, (T*)x, nparts_rz, part_rz, nparts_rr, part_rr, state)))
    } else {
        VT_SWITCH(dtype, NV_SWITCH(nvec, hipLaunchKernelGGL((cg_direction_fold_kernel<T, NV, false>), dim3(P), dim3(kKrylovThreads), 0, s, n, (const T*)z, (T*)p, (T*)x, nparts_rz, part_rz, nparts_rr, part_rr, state)))
    }
    return launch_status();
}

int wpc_cg_fold_grouped(int step, int ntasks, const void* tasks, int total_blocks, int dtype,
                        void* stream) {
    if (ntasks <= 0) return WPC_OK;
    if (tasks == nullptr || total_blocks <= 0 || (step != 1 && step != 2)) return WPC_EARG;
    hipStream_t s = as_stream(stream);
    const CgTask* t = (const CgTask*)tasks;
    // one instantiation for every vector count: each task carries its own
    if (step == 1) {
        VT_SWITCH(dtype, hipLaunchKernelGGL((grouped_update_fold_kernel<T, 0>), dim3(total_blocks), dim3(kKrylovThreads), 0, s, t, ntasks))
    } else {
        VT_SWITCH(dtype, hipLaunchKernelGGL((grouped_direction_fold_kernel<T, 0>), dim3(total_blocks), dim3(kKrylovThreads), 0, s, t, ntasks))
    }
    return launch_status();
}

const char* wpc_version(void) { return "wpc 0.1 gfx950"; }

}  // extern "C"
