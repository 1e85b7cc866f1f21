/* wpc.h -- C ABI of the MI355X (gfx950) warp-preconditioners hot path.
 *
 * Every entry point is stream-ordered and takes plain device pointers, element
 * counts and a HIP stream (`stream`, a hipStream_t passed as void*; NULL is the
 * null stream).  No entry point allocates device memory: the caller owns every
 * buffer, including the scratch (`workspace`, `tmp`, `partials`) arguments.
 * Indices are int32 (CSR offsets/columns, as in Warp's BsrMatrix).  Status
 * counters are int32 device arrays the kernels atomically increment; the caller
 * zeroes them and reads them back.
 *
 * Return value: WPC_OK (0), a negative WPC_E* code for an argument the library
 * does not support, or a positive hipError_t from the launch.
 *
 * Each declaration cites the reference routine it replaces (paths relative to
 * the root of alecjacobson/warp-preconditioners).  The reference is pure Python
 * over NVIDIA Warp; these are the kernels a ctypes layer binds in its place.
 */
#ifndef WPC_H
#define WPC_H

#ifdef __cplusplus
extern "C" {
#endif

/* Value / vector storage formats. */
enum {
    WPC_F16 = 1,  /* IEEE binary16 */
    WPC_BF16 = 2, /* bfloat16 (factor storage only) */
    WPC_F32 = 3,
    WPC_F64 = 4
};

enum {
    WPC_OK = 0,
    WPC_EDTYPE = -1,  /* unsupported storage/vector format combination */
    WPC_EARG = -2,    /* size, width, lanes or vector-count out of range */
    WPC_ESCRATCH = -3 /* workspace too small */
};

/* Library identity: returns a static string (build id and target). */
const char* wpc_version(void);

/* ------------------------------------------------------------ sparse products */

/* z = alpha * diag(row_scale) * A * x + beta * y for scalar CSR A.
 * Replaces sparse_operator.py:90-104 (_matvec: the grouped multi-lane kernel
 * :10-40, the mixed-storage kernel :64-87 and Warp's bsr_mv).
 * `lanes` (1,2,4,8,16,32,64; 0 = pick by mean row length) lanes cooperate on
 * each row.  `nvec` (1..8) vectors are stored interleaved: x[col*nvec+v].
 * row_scale (nullable, float32[n]) multiplies each row sum before alpha.
 * alpha == 0 never reads A or x; beta == 0 never reads y.  x must not alias z;
 * y may alias z.  Accumulation is in the vector format (float for binary16
 * vectors).  val_dtype may be narrower than vec_dtype (mixed storage). */
int wpc_csr_spmv(int n, const int* offsets, const int* columns, const void* values, int val_dtype,
                 const float* row_scale, const void* x, const void* y, void* z, int vec_dtype,
                 int nvec, double alpha, double beta, int lanes, void* stream);

/* z = alpha * A * x + beta * y for b x b BSR (b = 1..4), same format for values
 * and vectors (F32/F64).  Replaces Warp's bsr_mv on block matrices (CG on a
 * BsrMatrix operator; biharmonic.py:103,106).  x must not alias z. */
int wpc_bsr_spmv(int nb, int b, const int* offsets, const int* columns, const void* values,
                 int dtype, const void* x, const void* y, void* z, double alpha, double beta,
                 void* stream);

/* z = alpha * G^T (G x) + beta * y through tmp[n*nvec].  Replaces FSAI._apply
 * (fsai.py:374-381).  G is lower triangular CSR with its transpose GT stored
 * explicitly.  With row_scale2 (nullable, float32[n]) the stored factor is
 * q = G / s row-wise and the product is G^T G = q^T diag(s^2) q: the scale is
 * applied in the first pass, GT holds q^T unscaled.  x is fully consumed before
 * z is written, so x, y and z may alias.  alpha == 0 never reads x. */
int wpc_fsai_apply(int n, const int* g_offsets, const int* g_columns, const void* g_values,
                   const int* gt_offsets, const int* gt_columns, const void* gt_values,
                   int factor_dtype, const float* row_scale2, void* tmp, const void* x,
                   const void* y, void* z, int vec_dtype, int nvec, double alpha, double beta,
                   int lanes, void* stream);

/* wpc_fsai_apply with its temporary in a library-owned device array instead of
 * `tmp`: allocates nothing (no hipMalloc, hipFree or allocating library call).
 * No counterpart in the reference; for callers that cannot allocate device
 * memory (code under graph capture, serving loops, harnesses that forbid it).
 * n * nvec * sizeof(vector element) must not exceed WPC_NOALLOC_SCRATCH_BYTES
 * (else WPC_ESCRATCH, nothing launched).  Each GPU holds one such array until the
 * process ends, shared by every call on that GPU: calls that overlap on one GPU
 * race, so the caller serialises them.  Same kernels and results as
 * wpc_fsai_apply. */
#define WPC_NOALLOC_SCRATCH_BYTES (32u << 20)
int wpc_fsai_apply_noalloc(int n, const int* g_offsets, const int* g_columns,
                           const void* g_values, const int* gt_offsets, const int* gt_columns,
                           const void* gt_values, int factor_dtype, const float* row_scale2,
                           const void* x, const void* y, void* z, int vec_dtype, int nvec,
                           double alpha, double beta, int lanes, void* stream);

/* Row partition for wpc_csr_stream (host function, no device work): rows are cut
 * into blocks of at most 256 rows and 2048 stored entries, a longer row gets a
 * block of its own.  host_blocks receives count+1 row boundaries; returns count
 * (or WPC_ESCRATCH if capacity is too small). */
int wpc_csr_row_blocks(int n, const int* host_offsets, int* host_blocks, int capacity);

/* Sparse product with a fused epilogue, in one of two row mappings:
 *   lanes == 0  streamed: each workgroup loads its row block's (from `blocks`,
 *               see wpc_csr_row_blocks) column/value range into LDS with
 *               coalesced loads and one lane per row sums the row serially in
 *               storage order (the order of sparse_operator.py:64-87 and of
 *               Warp's CSR bsr_mv);
 *   lanes = 2^k (1..64) lanes own one row and stride over it (the grouped
 *               kernel of sparse_operator.py:10-40); blocks may be NULL.
 * Computes  z = alpha * (s .* (A x) + d .* u) + beta * y  with optional row
 * scale s (float32 or vec_dtype), optional diagonal term d .* u (d: vec_dtype[n],
 * u: like x); when w is non-null, also one partial per (workgroup, vector) of
 * sum_i w[i,v] z[i,v] into partials (double[workgroups * nvec]; workgroups =
 * nblocks when streamed, ceil(n*lanes/256) otherwise), folded in workgroup order
 * into out[v] by a single-workgroup launch when out is non-null.  alpha == 0
 * never reads A, x or u; beta == 0 never reads y; x must not alias z.  Replaces
 * sparse_operator.py:90-104 (_matvec), the two passes of fsai.py:374-381 and of
 * squared_laplacian.py:37-98 (forward/backward, identity free map). */
int wpc_csr_stream(int nblocks, const int* blocks, int n, int lanes, const int* offsets,
                   const int* columns, const void* values, int val_dtype, const void* s,
                   int s_dtype, const void* d, const void* u, const void* x, const void* y,
                   void* z, int vec_dtype, int nvec, double alpha, double beta, const void* w,
                   double* partials, double* out, void* stream);

/* Sliced-ELL copy of a scalar CSR matrix for the sliced product mapping: rows in
 * slices of 64, a slice's entries column-major (entry k of the row in position r
 * of slice s at (slice_ptr[s] + k) * 64 + r), rows padded to the slice's longest
 * row with column -1 and value 0.  Position i holds row rows[i] when `rows` (a
 * permutation of 0..n-1) is given, row i when it is NULL.  Build:
 * wpc_sell_lengths -> slice_ptr[0] = 0, slice_ptr[1..] = inclusive scan of
 * slice_len (wpc_scan_inclusive) -> allocate 64 * slice_ptr[nslices] entries ->
 * wpc_sell_fill -> wpc_sell_values (rerun after the CSR values change).  No
 * reference counterpart: a layout of this library for the products of
 * sparse_operator.py:14-104 / fsai.py:374-381. */
int wpc_sell_slices(int n);
int wpc_sell_lengths(int n, const int* offsets, const int* rows, int* slice_len, void* stream);
int wpc_sell_fill(int n, const int* offsets, const int* columns, const int* rows,
                  const int* slice_ptr, const int* slice_len, int* sell_columns, int* sell_map,
                  void* stream);
int wpc_sell_values(long count, const int* sell_map, const void* values, int elem_bytes,
                    void* sell_values, void* stream);
/* 16-bit copy of the sliced columns (run after wpc_sell_fill): slice_base[s]
 * (int32, one per slice) = the smallest column of slice s; each entry stores
 * column - slice_base[s] as an unsigned 16-bit offset (in the bits of
 * `sell_col16`), padding 0xFFFF.  overflow[0] (zeroed by the caller) counts
 * slices whose columns span more than 0xFFFE, in which case the int32 columns
 * must be used.  Halves the index stream of banded (mesh-ordered) matrices. */
int wpc_sell_col16(int n, const int* offsets, const int* rows, const int* slice_ptr,
                   const int* slice_len, const int* sell_columns, short* sell_col16,
                   int* slice_base, int* overflow, void* stream);
/* Stencil form of the sliced copy (natural row order, no column stream): the
 * offsets column - row of a slice's rows, merged in ascending order, are the
 * slice's entries.  Entry k of the row in position r of slice s has column
 * r + deltas[32 * s + k] where bit k of row_mask[r] is set; a clear bit is
 * padding (value 0).  Each row's entries stay in the order of its columns, so the
 * products' sums equal the plain slices' (up to the sign of a zero sum).  Run
 * twice: with slice_ptr = NULL it writes slice_len[s] = the number of distinct
 * offsets of slice s and adds to status[0] (zeroed by the caller) the slices with
 * more than 32 (then the form does not apply); after slice_ptr[0] = 0, slice_ptr[1..]
 * = inclusive scan of slice_len, it writes deltas (32 * nslices ints), row_mask
 * (n ints) and sell_map (64 * slice_ptr[nslices] ints, -1 for padding) for
 * wpc_sell_values.  The products take it as col16 = 2 with sell_columns = deltas
 * and slice_base = row_mask (ncol * nvec * sizeof(vector) must not exceed 2^31).
 * No reference counterpart: a layout of this library for mesh-ordered matrices,
 * whose slices share a few offsets (the 15 of a Kuhn grid's P1 Laplacian). */
int wpc_sell_stencil(int n, const int* offsets, const int* columns, const int* slice_ptr,
                     int* slice_len, int* deltas, int* row_mask, int* sell_map, int* status,
                     void* stream);

/* The product of wpc_csr_stream on the sliced copy: one thread per row, storage
 * order serial sums (bit-identical to the streamed mapping), same epilogue and
 * fused inner product; rows as given to wpc_sell_fill; sell_columns are the int32
 * columns (col16 = 0; then ncol * nvec * sizeof(vector) must not exceed 2^31), the
 * 16-bit offsets of wpc_sell_col16 (col16 = 1) or the deltas of wpc_sell_stencil
 * (col16 = 2, rows NULL, slice_base = its row masks; range as for int32).
 * slice_base (one int32 per slice) is always read: the bases of wpc_sell_col16, or
 * zeros with int32 columns.
 * Padding adds exactly +0 whatever x holds (its gathers fall outside the buffer
 * descriptor's range and read 0).  partials: ceil(n / 256) workgroups (workgroup b
 * covers positions 256b..+255). */
int wpc_sell_stream(int n, const int* slice_ptr, const int* slice_len, const int* slice_base,
                    const int* rows, const void* sell_columns, int col16,
                    const void* sell_values, int val_dtype,
                    const void* s, int s_dtype, const void* d, const void* u, const void* x,
                    const void* y, void* z, int vec_dtype, int nvec, double alpha, double beta,
                    const void* w, double* partials, double* out, void* stream);

/* 3x3 BSR product on a sliced block copy: z = alpha * A x + beta * y, vectors
 * interleaved [scalar row][nvec] (nvec 1..4; scalar row 3i + c of block row i).
 * The block pattern's sliced copy comes from wpc_sell_lengths / wpc_sell_fill /
 * wpc_sell_col16 run on the block offsets and columns (nb block rows; int32 block
 * columns or 16-bit offsets from slice_base); wpc_bsr3_sell_values stores each
 * slot's nine values component-major inside the slice (value q of the block in
 * slot p at ((p / 64) * 9 + q) * 64 + p % 64; `slots` = 64 * slice_ptr[nslices],
 * sell_map from wpc_sell_fill; rerun after the BSR values change).  One thread
 * per block row sums its blocks in storage order, each block's product
 * t_r = (v_r0 x_0 + v_r1 x_1) + v_r2 x_2 added to the row's sum, in the vector
 * type (values F64 or F32 with F64 vectors, F32 with F32).  With w: per-workgroup
 * partials of w . z (ceil(nb / 256) per vector), folded into out[nvec] when out
 * is given.  alpha == 0 never reads x, beta == 0 never reads y; x must not alias
 * z.  Replaces Warp's bsr_mv on 3x3 blocks (sparse_operator.py:104, the CG
 * operator of benchmarks/block_challenge.py:155,197 and benchmarks/elasticity.py). */
int wpc_bsr3_sell_values(long slots, const int* sell_map, const void* values, int elem_bytes,
                         void* sell_values, void* stream);
int wpc_bsr3_sell_stream(int nb, const int* slice_ptr, const int* slice_len,
                         const int* slice_base, const void* sell_columns, int col16,
                         const void* sell_values, int val_dtype, const void* x, const void* y,
                         void* z, int vec_dtype, int nvec, double alpha, double beta,
                         const void* w, double* partials, double* out, void* stream);

/* ------------------------------------------------------------ FSAI setup */

/* scale[i] = 1/sqrt(A_ii); non-positive or non-finite diagonals set 0 and add 1
 * to status[0].  Replaces fsai.py:43-57 (diagonal). dtype F16/F32/F64. */
int wpc_fsai_diagonal(int n, const int* offsets, const int* columns, const void* values, int dtype,
                      void* scale, int* status, void* stream);

/* Adaptive equilibrated FSAI rows, one wavefront per row.  Replaces fsai.py:59-176
 * (build).  Writes cols_out/vals_out[i*width+p] (-1/0 padded, selection order,
 * diagonal first), status[1] += truncated rows, status[2] += invalid rows.
 * width, step in [1,64]; tolerance is kap_tolerance, pivot_floor the local pivot
 * floor.  Arithmetic is that of `dtype` (F16 as IEEE binary16 operations). */
int wpc_fsai_build(int n, const int* offsets, const int* columns, const void* values, int dtype,
                   const void* scale, double tolerance, double pivot_floor, int width, int step,
                   int* cols_out, void* vals_out, int* status, void* stream);

/* Sort each padded row by column and drop -1 slots; counts[i+1] = row size
 * (counts[0] is set to 0).  Replaces fsai_pack.py:15-39 (sort_rows). */
int wpc_fsai_sort_rows(int n, int width, int* columns, void* values, int dtype, int* counts,
                       void* stream);

/* out[i] = in[0] + ... + in[i] (int32) in one workgroup.  Replaces
 * wp.utils.array_scan (fsai_pack.py:63). */
int wpc_scan_inclusive(int count, const int* in, int* out, void* stream);

/* Compact sorted padded rows into CSR.  Replaces fsai_pack.py:41-53 (pack). */
int wpc_fsai_pack(int n, int width, const int* offsets, const int* columns, const void* values,
                  int dtype, int* out_columns, void* out_values, void* stream);

/* Expand compact b x b BSR into scalar CSR (row r*b+c), keeping stored zeros.
 * Replaces fsai_pack.py:74-130 (scalarize / scalarize_compact). */
int wpc_bsr_scalarize(int nb, int b, const int* offsets, const int* columns, const void* values,
                      int dtype, int* out_offsets, int* out_columns, void* out_values,
                      void* stream);

/* Transpose scalar CSR (nrow x ncol, nnz = offsets[nrow]) with ascending
 * columns in every output row; values move as raw `elem_bytes` elements.
 * workspace: int32[ncol + 1].  Replaces Warp's bsr_transposed (fsai.py:337). */
int wpc_csr_transpose(int nrow, int ncol, int nnz, const int* offsets, const int* columns,
                      const void* values, int elem_bytes, int* t_offsets, int* t_columns,
                      void* t_values, int* workspace, void* stream);

/* dst[i] = (dst format) src[i].  Replaces bsr_copy(..., scalar_type=...) on
 * values (fsai.py:325) and wp.utils.array_cast. */
int wpc_convert(long count, const void* src, int src_dtype, void* dst, int dst_dtype,
                void* stream);

/* status[0] += rows with a non-finite value or a non-positive diagonal.
 * Replaces fsai.py:181-196 (_validate_float32_factor). */
int wpc_fsai_validate_factor(int n, const int* offsets, const int* columns, const void* values,
                             int dtype, int* status, void* stream);

/* Row-scaled narrow factor: s_i = max_j |G_ij|, q = G / s_i stored in q_dtype
 * (F16/BF16/F32), row_scale2[i] = s_i^2 (float32).  status[0] += rows with a
 * non-finite value or a non-positive stored diagonal (diagonal last).  No
 * counterpart in the reference (its float16 factors are unscaled); it feeds
 * wpc_fsai_apply's row_scale2 path. */
int wpc_fsai_compress_scaled(int n, const int* offsets, const int* columns, const void* values,
                             int dtype, void* q_values, int q_dtype, float* row_scale2,
                             int* status, void* stream);

/* ------------------------------------------------------------ FSAI refit */

/* bad[0] += number of positions where a != b.  Replaces fsai_reuse.py:10-14. */
int wpc_compare_i32(long count, const int* a, const int* b, int* bad, void* stream);

/* Scalar nz -> index into the source BSR value array.  Replaces fsai_reuse.py:35-48. */
int wpc_refit_source_map(int n, const int* offsets, const int* columns, const int* src_offsets,
                         const int* src_columns, int block, int* mapping, void* stream);

/* G nz -> position in GT.  Replaces fsai_reuse.py:51-61. */
int wpc_refit_transpose_map(int n, const int* offsets, const int* columns, const int* t_offsets,
                            const int* t_columns, int* mapping, void* stream);

/* out[e] = src[map[e]] / out[map[e]] = src[e] on `elem_bytes` elements.
 * Replace fsai_reuse.py:69-74 (gather) and :136-141 (transpose). */
int wpc_gather(long count, int elem_bytes, const void* src, const int* map, void* out,
               void* stream);
int wpc_scatter(long count, int elem_bytes, const void* src, const int* map, void* out,
                void* stream);

/* Fixed-support refit: per row, dense Cholesky of the equilibrated principal
 * submatrix on G's stored (sorted, diagonal last) support and
 * G[i,S] = z/sqrt(z_last)*scale[S] with z = A[S,S]^-1 e_last.  Writes
 * out[offsets of G] in out_dtype; bad[0] += invalid rows.  Replaces
 * fsai_reuse.py:76-134 (refit). */
int wpc_fsai_refit(int n, const int* offsets, const int* columns, const void* values, int dtype,
                   const int* g_offsets, const int* g_columns, const void* scale,
                   double pivot_floor, int width, void* out, int out_dtype, int* bad,
                   void* stream);

/* out[i] = +-1 (Rademacher), deterministic in (seed, i).  Stands in for
 * fsai_reuse.py:143-147 (probe); the random stream differs from Warp's. */
int wpc_rademacher(int n, int seed, void* out, int dtype, void* stream);

/* ------------------------------------------------------------ vectors */

/* z = alpha * x + beta * y elementwise (count elements; alpha == 0 never reads
 * x, beta == 0 never reads y).  Replaces fsai_reuse.py:149-152 (subtract),
 * block_ilu.py:173-185 (combine) and Warp's array_axpy. */
int wpc_axpby(long count, double alpha, const void* x, double beta, const void* y, void* z,
              int dtype, void* stream);

/* out[v] = sum_i x[i*nvec+v]*y[i*nvec+v] in float64, deterministic order
 * (per-block partials, then one fixed-order pass); `rows` = count / nvec.
 * partials: double[wpc_dot_partials(rows) * nvec].  Replaces
 * wp.utils.array_inner (fsai_reuse.py:278). */
int wpc_dot_partials(long rows);
int wpc_dot(long rows, int nvec, const void* x, const void* y, int dtype, double* partials,
            double* out, void* stream);

/* dst[i] = value (count elements of dtype) / int32 fill. */
int wpc_fill(long count, void* dst, int dtype, double value, void* stream);
int wpc_fill_i32(long count, int* dst, int value, void* stream);

/* ------------------------------------------------------------ squared Laplacian */

/* index[free[i]] = i with duplicate/range detection (index pre-filled with -1).
 * status[0] += out-of-range, status[1] += duplicates.  Replaces
 * squared_laplacian.py:11-20 (_index_free). */
int wpc_sqlap_index_free(int n_free, const int* free_indices, int n, int* index, int* status,
                         void* stream);

/* inverse[i] = 1/mass[i]; status[2] += non-positive/non-finite masses.
 * Replaces squared_laplacian.py:27-35 (mass_inverse). */
int wpc_sqlap_mass_inverse(int n, const void* mass, void* inverse, int dtype, int* status,
                           void* stream);

/* Factored (L M^-1 L + w M)[f,f] product.  Replaces squared_laplacian.py:37-98
 * (forward/backward) as driven by _product :437-474.
 *   pass 1: tmp[r] = inverse[r] * sum_{e in row r} L_e * x~[col]   (all rows of L)
 *           x~[col] = x[index[col]] if index[col] >= 0 (boundary = 0), or
 *                     x[col] if index[col] < 0 (boundary = 1)
 *   pass 2: z[i] = alpha*(sum_{e in row free[i]} L_e tmp[col] + w*mass[free[i]]*x[i])
 *                  + beta*y[i]
 * index/free may be NULL when every row is free (identity maps).  nvec
 * interleaved vectors.  x, y, z may alias (x is consumed by pass 1 except for
 * the same-row mass term).  tmp: n_full*nvec. */
int wpc_sqlap_apply(int n_full, const int* offsets, const int* columns, const void* values,
                    int dtype, const void* inverse, const int* index, int n_free,
                    const int* free_indices, const void* mass, double mass_weight, void* tmp,
                    const void* x, const void* y, void* z, int nvec, double alpha, double beta,
                    int boundary, int lanes, void* stream);

/* Block-adaptive FSAI on b x b BSR matrices (b = 2..4, b * max_blocks <= 32).
 * Replaces benchmarks/block_fsai_candidate.py:45-176 (diagonal + build); with
 * refit != 0 the block pattern is read back from `cols` (:76-82, the rows as
 * sorted by wpc_fsai_sort_rows).  Writes scale[n], padded cols/vals[n * b *
 * max_blocks] in pattern order; status[0] counts nonpositive pivots. */
int wpc_block_fsai_build(int nb, int b, int max_blocks, const int* offsets, const int* columns,
                         const void* values, int dtype, void* scale, int* cols, void* vals,
                         int* status, int refit, void* stream);

/* ------------------------------------------------------------ assembly */

/* P1 tetrahedra: local[16*t + 4a + b] = vol_t * grad(phi_a) . grad(phi_b) and
 * quarter_volume[t] = vol_t / 4 (vertices float64 [nv,3], tets int32 [ntet,4]).
 * Replaces the host stiffness/mass assembly the reference's benchmarks run
 * (benchmarks/elasticity_problem.py:33-61, scalar term; lumped mass :44). */
int wpc_p1_local(int ntet, const double* vertices, const int* tets, double* local,
                 double* quarter_volume, void* stream);

/* out[s] = sum_{k = seg[s]}^{seg[s+1]-1} values[order[k]] summed in k order
 * (order may be NULL = identity): deterministic triplet assembly after a stable
 * key sort (Warp's bsr_from_triplets duplicate summation, biharmonic.py:97,169). */
int wpc_segment_sum(int nseg, const long* segments, const long* order, const double* values,
                    double* out, void* stream);

/* Q = K diag(1/m) K + alpha diag(m) on the device (float64): per row, the products
 * (K[i,k] * inverse_mass[k]) * K[k,j] accumulate per column in the order of k
 * along row i, then alpha*m_i joins the diagonal -- the term order of the host
 * SpGEMM the reference uses for its assembled squared operator
 * (visualization/solve_dirichlet.py:150-153).  count: counts[i] = row size of Q
 * (-1 if a row exceeds 512 distinct columns); scan to q_offsets, then fill
 * (columns sorted). */
int wpc_spgemm_kmk_count(int n, const int* offsets, const int* columns, int* counts,
                         void* stream);
int wpc_spgemm_kmk_fill(int n, const int* offsets, const int* columns, const double* values,
                        const double* inverse_mass, const double* mass, double alpha,
                        const int* q_offsets, int* q_columns, double* q_values, void* stream);

/* ------------------------------------------------------------ block Jacobi / ILU(0) */

/* rows[e] = block row of e; diagonal[i] = slot of block (i,i) or -1 (bad[0]++).
 * Replaces block_ilu.py:45-61 (topology). */
int wpc_block_topology(int nb, const int* offsets, const int* columns, int* rows, int* diagonal,
                       int* bad, void* stream);

/* inverse[i] = safe_inverse(values[diagonal[i]]) for b = 1..4.  Replaces
 * block_ilu.py:16-43 (safe_inverse) + :63-73 (invert). */
int wpc_block_invert(int nb, int b, const void* values, const int* diagonal, void* inverse,
                     double tolerance, int* bad, int dtype, void* stream);

/* Binary-search each diagonal block and invert it.  Replaces block_ilu.py:75-104
 * (diagonal_only; compact storage). */
int wpc_block_diagonal_only(int nb, int b, const int* offsets, const int* columns,
                            const void* values, void* inverse, double tolerance, int* bad,
                            int dtype, void* stream);

/* factors[e] = values[e] * inverse[col] below the diagonal, else values[e].
 * Replaces block_ilu.py:106-118 (initialize). */
int wpc_ilu_initialize(long nnzb, int b, const int* rows, const int* columns, const void* values,
                       const void* inverse, void* factors, int dtype, void* stream);

/* One synchronous ILU(0) sweep with relaxation.  Replaces block_ilu.py:120-150. */
int wpc_ilu_sweep(long nnzb, int b, const int* offsets, const int* columns, const int* rows,
                  const void* values, const void* old, const void* inverse, void* out,
                  double relaxation, int dtype, void* stream);

/* One Jacobi step of the unit-lower (lower=1) or upper (lower=0) solve.
 * Replaces block_ilu.py:152-171 (triangular). */
int wpc_ilu_triangular(int nb, int b, const int* offsets, const int* columns, const void* factors,
                       const void* inverse, const void* rhs, const void* old, void* out,
                       int lower, int dtype, void* stream);

/* z_i = alpha * D_i^-1 x_i + beta * y_i.  Replaces block_ilu.py:187-202. */
int wpc_block_diagonal_apply(int nb, int b, const void* inverse, const void* x, const void* y,
                             void* z, double alpha, double beta, int dtype, void* stream);

/* bad[0] += non-finite entries.  Replaces block_ilu.py:204-210 (validate_factors). */
int wpc_count_nonfinite(long count, const void* values, int dtype, int* bad, void* stream);

/* bad[0] += entries that are not finite and positive.  Replaces biharmonic.py:70-74. */
int wpc_count_nonpositive(long count, const void* values, int dtype, int* bad, void* stream);

/* ------------------------------------------------------------ Krylov (batched) */

/* Vector steps of warp.optim.linear.cg / cr (the reference's solvers, called e.g.
 * at tests/test_fsai.py:64 and benchmarks/synthetic.py:50), for nvec right-hand
 * sides on interleaved [n][nvec] vectors in dtype F32 or F64.  Scalars live on
 * the device in a float64 array state[WPC_KSTATE * nvec], field-major
 * (state[field * nvec + v]).  A slot whose incoming r.r already meets its
 * target takes no step (device-side stopping), so a captured chunk of
 * iterations may run past convergence without changing the result; ITER counts
 * the steps actually taken.
 *
 * One CG iteration:  Ap = A p;  krylov_dot(p, Ap -> pd);  krylov_update;
 *                    z = M r;   krylov_dot(r, z -> pz);  krylov_direction.
 * One CR iteration:  MAp = M Ap; krylov_dot(Ap, MAp -> pd); krylov_update (also
 *                    z -= alpha MAp);  Az = A z; krylov_dot(z, Az -> pz);
 *                    krylov_direction (also Ap = Az + beta Ap). */
enum {
    WPC_KS_RZ = 0,      /* r.z (CG) or z.Az (CR) of the current direction */
    WPC_KS_RR = 1,      /* r.r */
    WPC_KS_TARGET = 2,  /* squared stopping threshold */
    WPC_KS_ALPHA = 3,
    WPC_KS_ITER = 4,    /* steps taken */
    WPC_KS_DENOM = 5,   /* p.Ap (CG) or Ap.MAp (CR) */
    WPC_KS_RZ_OLD = 6,
    WPC_KS_ACTIVE = 7,  /* the current step moves this slot */
    WPC_KS_RR_NEW = 8,  /* r.r after the update (fused CG) */
    WPC_KS_RZ_NEW = 9,  /* r.z of the new preconditioned residual (fused CG) */
    WPC_KS_BETA = 10,
    WPC_KSTATE = 12
};

/* Number of blocks P the Krylov kernels use for n rows; each partials region is
 * double[P * nvec]. */
int wpc_krylov_partials(int n);

/* partials[b*nvec+v] = block b's share of sum_i a[i,v]*b[i,v]. */
int wpc_krylov_dot(int n, int nvec, const void* a, const void* b, int dtype, double* partials,
                   void* stream);

/* Folds part_denom; alpha = active ? rz/denom : 0; x += alpha p; r -= alpha Ap
 * (CR: z -= alpha map); writes r.r partials to part_rr.  method 0 = CG, 1 = CR. */
int wpc_krylov_update(int method, int n, int nvec, void* x, void* r, const void* p,
                      const void* ap, void* z, const void* map, int dtype, double* state,
                      const double* part_denom, double* part_rr, void* stream);

/* Folds part_rz and part_rr; for slots that stepped: beta = rz_new/rz_old,
 * p = z + beta p (CR: ap = az + beta ap), state RZ/RR/ITER updated. */
int wpc_krylov_direction(int method, int n, int nvec, const void* z, void* p, const void* az,
                         void* ap, int dtype, double* state, const double* part_rz,
                         const double* part_rr, void* stream);

/* state[TARGET] holds ||b||^2 on entry; becomes max(atol, tol*||b||)^2; ITER = 0. */
int wpc_krylov_set_target(int nvec, double* state, double tol, double atol, void* stream);

/* Fused CG (no host round trip; every reduction folded in a fixed order):
 *   A p              operator whose last pass leaves p.Ap partials
 *                    (wpc_csr_stream with w = p, out = NULL)
 *   wpc_cg_alpha     DENOM = fold, ACTIVE = RR > TARGET, ALPHA = RZ/DENOM or 0
 *   wpc_cg_update    x += ALPHA p, r -= ALPHA Ap, r.r partials
 *   M r              preconditioner whose last pass leaves r.z partials
 *   wpc_cg_beta      RZ_NEW, RR_NEW folded; stepping slots: BETA = RZ_NEW/RZ,
 *                    RZ <- RZ_NEW, RR <- RR_NEW, ITER += 1
 *   wpc_cg_direction p = z + BETA p for stepping slots.
 * update partials: double[wpc_krylov_partials(n) * nvec]. */
int wpc_cg_alpha(int nvec, int nparts, const double* part_denom, double* state, void* stream);
int wpc_cg_update(int n, int nvec, void* x, void* r, const void* p, const void* ap, int dtype,
                  const double* state, double* partials, void* stream);
int wpc_cg_beta(int nvec, int nparts_rz, const double* part_rz, int nparts_rr,
                const double* part_rr, double* state, void* stream);
int wpc_cg_direction(int n, int nvec, const void* z, void* p, int dtype, const double* state,
                     void* stream);

/* The same iteration in four launches instead of six: every workgroup of the
 * update folds the p.Ap partials (and of the direction step the r.z and r.r
 * partials) in one fixed order, so all workgroups agree on ALPHA/BETA bit for bit
 * and workgroup 0 records the scalars.  Meant for small partial counts
 * (nparts * nvec <= 4096: the folds are repeated per workgroup).  The pair is used
 * together: the update saves RZ in RZ_OLD, which the direction step reads.
 * x = NULL in the update leaves x (and p) alone; the direction step given x then
 * applies the update's x += ALPHA p before it replaces p (one vector read fewer per
 * iteration, the same arithmetic).  Pass x to exactly one of the two. */
int wpc_cg_update_fold(int n, int nvec, void* x, void* r, const void* p, const void* ap, int dtype,
                       int nparts, const double* part_denom, double* state, double* partials,
                       void* stream);
int wpc_cg_direction_fold(int n, int nvec, const void* z, void* p, void* x, int dtype,
                          int nparts_rz, const double* part_rz, int nparts_rr,
                          const double* part_rr, double* state, void* stream);

/* The factored squared-Laplacian operator's two sliced passes inside fused CG, with
 * the update step folded into the second pass: replaces, for an operator with the
 * identity free map, the product pair (squared_laplacian.py:37-65 forward,
 * :67-98 backward, applied by Warp's cg as A p) followed by wpc_cg_update_fold
 * with x = NULL.  Pass 1: t = inverse_mass .* (L p) and per-workgroup partials of
 * the curvature p.Ap = sum_i (L p)_i t_i + W_i p_i^2 (W = weighted_mass = w * mass,
 * NULL for w = 0; L symmetric).  Pass 2: Ap = L t + W p row by row (not stored),
 * ALPHA = RZ / (fold of the pass-1 partials) for slots with RR > TARGET (else 0),
 * r -= ALPHA Ap, per-workgroup partials of r.r; workgroup 0 writes RZ_OLD, DENOM,
 * ACTIVE, ALPHA as wpc_cg_update_fold does; with x non-NULL pass 2 also advances
 * x += ALPHA p (it reads p anyway) and the direction step folds the
 * wpc_sqlap_cg_partials(n) r.r partials from part_rr with x = NULL (else
 * wpc_cg_direction_fold advances x).  The sliced copy of L (rows in natural or given
 * order) as for wpc_sell_stream; values, inverse_mass, weighted_mass and vectors
 * all F64 or all F32; nvec 1..8; part_energy and part_rr hold
 * wpc_sqlap_cg_partials(n) * nvec doubles each (the energy partials are folded in
 * place by every workgroup of pass 2: meant for counts times nvec <= 4096).
 * The curvature differs from p . (A p) by rounding only.  passes = 3 runs both;
 * 1 or 2 runs one of them (for timing each kernel alone).  Bit 2 (passes 6 or 7),
 * for partial counts too large to fold in every workgroup: pass 2 takes ALPHA as
 * wpc_cg_alpha left it in the state after folding the pass-1 partials (no fold,
 * no state writes) and advances x += ALPHA p (x required), as wpc_cg_update does;
 * the step then ends with wpc_cg_beta (these r.r partials) and wpc_cg_direction.
 * x, when given, must not alias p, t or r. */
int wpc_sqlap_cg_update(int n, const int* slice_ptr, const int* slice_len,
                        const int* slice_base, const int* rows, const void* sell_columns,
                        int col16, const void* sell_values, int val_dtype,
                        const void* inverse_mass, const void* weighted_mass, const void* p,
                        void* t, void* r, int vec_dtype, int nvec, double* part_energy,
                        double* part_rr, double* state, int passes, void* x, void* stream);

/* Partials per vector that each pass of wpc_sqlap_cg_update leaves (one per
 * workgroup: ceil(n / 512)).  Host-only helper; no reference counterpart. */
int wpc_sqlap_cg_partials(int n);

/* No reference counterpart (warp.optim.linear.cg, the reference's solver, runs its
 * direction step on its own): the first pass of wpc_sqlap_cg_update with the
 * previous step's direction step folded in, for the steps of a CG run between the
 * first and the last, whose direction step it replaces.  Folds nparts_rz r.z
 * partials (part_rz, the preconditioner's last pass) and the
 * wpc_sqlap_cg_partials(n) r.r partials that pass 2 left (part_rr) into BETA and
 * writes the state as wpc_cg_direction_fold does (x is advanced by pass 2, not
 * here); then, per stepping slot, p = z + BETA p and t = inverse_mass .* (L z) +
 * BETA t, where t is what the previous first pass (either form) left; resting
 * slots keep p and t.  Leaves pass 2's curvature partials in part_energy, from
 * (L p)_i = t_i / inverse_mass_i.  Pass 2 (wpc_sqlap_cg_update, passes = 2) then
 * reads this t.  t is the carried vector and differs from inverse_mass .* (L p)
 * by the recurrence's rounding; z, p and t must not alias.  With part_rz = NULL
 * (partial counts too large to fold in every workgroup; nparts_rz and part_rr
 * ignored) BETA and ACTIVE are read from the state as wpc_cg_beta left them:
 * nothing is folded and the state is not written, and pass 2 follows as
 * wpc_cg_alpha + passes = 6.  Other arguments as for wpc_sqlap_cg_update. */
int wpc_sqlap_cg_direction_energy(int n, const int* slice_ptr, const int* slice_len,
                                  const int* slice_base, const int* rows,
                                  const void* sell_columns, int col16, const void* sell_values,
                                  int val_dtype, const void* inverse_mass,
                                  const void* weighted_mass, const void* z, void* p, void* t,
                                  int vec_dtype, int nvec, int nparts_rz, const double* part_rz,
                                  const double* part_rr, double* part_energy, double* state,
                                  void* stream);

/* ------------------------------------------------------------------ grouped CG steps
 * No reference counterpart: a serving extension that advances the fused CG of
 * several systems (uneven sizes and right-hand-side counts, same vector format
 * and product flags) with one launch per pass instead of one per system.  `tasks`
 * is a device table of 16 int64 words per system; systems own consecutive global workgroups
 * starting at their block0 (ascending, the first 0), total_blocks in all; tasks
 * with n = 0 (block0 = the end of the real ones) pad a launch captured for a
 * larger group and do nothing.  Each system computes exactly what its own
 * launch computes, bit for bit.
 *
 * wpc_sell_stream_grouped: per task {slice_ptr, slice_len, rows, sell_columns,
 *   sell_values, s, d, u, x, z, w, partials, n, block0, slice_base,
 *   nvec | 256 * col16} -- the wpc_sell_stream
 *   product with beta = 0 on each system (ceil(n/256) workgroups); s_dtype 0 means
 *   no row scale, has_d selects the d .* u term, dot the w . z partials.
 * wpc_cg_fold_grouped: step 1 = wpc_cg_update_fold with x = NULL, step 2 =
 *   wpc_cg_direction_fold with x, per task {x, r, p, ap, z, part_a, part_b, state,
 *   partials, n, nblocks, block0, pa, pb, nvec, 0} with nblocks =
 *   wpc_krylov_partials(n); update folds pa partials of p.Ap from part_a and leaves
 *   r.r partials in `partials`; direction folds pa r.z partials (part_a) and pb r.r
 *   partials (part_b) and advances x. */
int wpc_sell_stream_grouped(int ntasks, const void* tasks, int total_blocks, int val_dtype,
                            int s_dtype, int has_d, int dot, int vec_dtype, double alpha,
                            void* stream);
int wpc_cg_fold_grouped(int step, int ntasks, const void* tasks, int total_blocks, int dtype,
                        void* stream);

/* ------------------------------------------------------------------ certified residuals
 * No single reference routine: these supply the float64 true residual b - A x
 * that the reference's solves check through Warp's cg/cr (`bsr_mv` at
 * sparse_operator.py:104, squared_laplacian.py:158-195), accumulated with the
 * compensated Dot2 product so that the residual itself carries no cancellation
 * noise.  Vectors are float64 [n][nvec] (nvec 1..8).
 *   assembled operator (t_hi = t_lo = NULL):  r = b - A x        (values = A)
 *   factored operator:  t = inverse_mass .* (K x) kept as t_hi + t_lo, then
 *                       r = b - alpha*mass .* x - K t            (values = K)
 * norms[0..nvec) = sum r^2/mass, norms[nvec..2nvec) = sum mass*x^2 (fixed-order
 * fold of double[wpc_residual_partials(n) * 2 * nvec] block partials). */
int wpc_residual_partials(int n);
int wpc_dd_residual(int n, const int* offsets, const int* columns, const double* values,
                    const double* mass, const double* inverse_mass, double alpha, const double* x,
                    const double* b, double* t_hi, double* t_lo, double* r, int nvec,
                    double* partials, double* norms, void* stream);

/* Iterative-refinement steps, per right-hand side v with bit v of mask set
 * (s_v = sqrt(norms[v]), or 1 when zero):
 *   prepare: out = r / s_v (others 0), stored as out_dtype (f32 or f64)
 *   update:  d = s_v * c (others 0); x += d; c_out = d when c_out is not NULL
 * (c_out lets the caller certify x + d with wpc_dd_residual(b = r, x = c_out)).
 * b and r of wpc_dd_residual may be the same array. */
int wpc_refine_prepare(int n, int nvec, const double* r, const double* norms, int mask, void* out,
                       int out_dtype, void* stream);
int wpc_refine_update(int n, int nvec, double* x, const void* c, int c_dtype, const double* norms,
                      int mask, double* c_out, void* stream);

/* values[diag(i)] += alpha * d[i] on a canonical CSR matrix (forms K + alpha*M);
 * status[0] counts rows without a stored diagonal. */
int wpc_add_diagonal(int n, const int* offsets, const int* columns, double* values,
                     const double* d, double alpha, int* status, void* stream);

/* ------------------------------------------------------------------ GMRES basis
 * The Arnoldi products of restarted GMRES (warp.optim.linear.gmres as called at
 * tests/test_indefinite.py:130,218 and tests/test_biharmonic.py:32) on a row-major
 * basis V (row j at V + j*ld, ld >= n), 1 <= k <= 128 rows:
 *   wpc_gram:          h[j] = V[j] . w (double, deterministic two-stage sum;
 *                      partials: double[wpc_gram_partials(n) * k])
 *   wpc_basis_combine: w += alpha * sum_j h[j] V[j] (accumulated in double). */
int wpc_gram_partials(long n);
int wpc_gram(long n, int k, const void* V, long ld, const void* w, int dtype, double* partials,
             double* h, void* stream);
int wpc_basis_combine(long n, int k, const void* V, long ld, const double* h, double alpha,
                      void* w, int dtype, void* stream);

/* ------------------------------------------------------------------ mixed systems
 * Interleaved 2/3-field indefinite harmonic systems and their two block
 * preconditioners.  Vertex i of a system vector holds x[i*order + c]; fields
 * are n-vectors.  Arithmetic follows the reference expressions without
 * contraction.  dtype is WPC_F32 or WPC_F64 for every array of one call. */

/* mixed.py:19-57 (MixedHarmonicSystem's build kernel).  Writes n + nnz(K)
 * triplets: rows/cols int32, blocks [count][order][order], hvalues [count]
 * (the diagonal triplets carry weight*M / -M and shift*M, the others -K and
 * K); the caller sums duplicates.  bad[0] counts non-positive or non-finite
 * mass entries. */
int wpc_mixed_build(int n, int order, const int* offsets, const int* columns, const void* values,
                    const void* mass, double weight, double shift, int* rows, int* cols,
                    void* blocks, void* hvalues, int* bad, int dtype, void* stream);
/* mixed.py:66-79: out = x[:, c], or x[:, c] / mass when mass is not NULL. */
int wpc_mixed_component(int n, int order, const void* x, int c, const void* mass, void* out,
                        int dtype, void* stream);
/* mixed.py:81-86, 118-121: out = mass * x (divide = 0) or x / mass (divide = 1);
 * out may alias x. */
int wpc_mass_multiply(int n, const void* x, const void* mass, int divide, void* out, int dtype,
                      void* stream);
/* mixed.py:88-102: out = x[:, 0] - k1 (+ k2 when order = 3). */
int wpc_mixed_rhs(int n, int order, const void* x, const void* k1, const void* k2, void* out,
                  int dtype, void* stream);
/* mixed.py:104-116: out = -(ku + x[:, c]) / mass. */
int wpc_mixed_auxiliary(int n, int order, const void* x, int c, const void* ku, const void* mass,
                        void* out, int dtype, void* stream);
/* mixed.py:123-143: z = alpha * (u, a[, lam]) + beta * y; alpha = 0 reads no
 * field and beta = 0 never reads y; y and z may alias. */
int wpc_mixed_finish(int n, int order, const void* u, const void* a, const void* lam,
                     const void* y, void* z, double alpha, double beta, int dtype, void* stream);
/* mixed.py:330-345 (ShiftedBlock unpack); c is unused when order = 2. */
int wpc_shifted_unpack(int n, int order, const void* x, void* a, void* b, void* c, double scale,
                       int dtype, void* stream);
/* mixed.py:347-372 (ShiftedBlock finish): order 2 gives (a / scale, b * scale),
 * order 3 (a, b, c), then z = alpha * value + beta * y as above. */
int wpc_shifted_finish(int n, int order, const void* a, const void* b, const void* c,
                       const void* y, void* z, double scale, double alpha, double beta, int dtype,
                       void* stream);
/* mixed.py:147-160: out[k] = s_{k mod order} * x[k] for count = n * order
 * interleaved entries (order 1 scales a plain vector by s0); out may alias x. */
int wpc_vector_scale(long count, int order, const void* x, void* out, double s0, double s1,
                     double s2, int dtype, void* stream);

/* ------------------------------------------------------------------ graphs
 * No reference counterpart (Warp captures with its own graph API around
 * cg/cr): capture the library launches issued on `stream` between begin and end
 * into an executable HIP graph (thread-local capture mode, no device
 * synchronization); launch it on any stream; destroy it when done. */
int wpc_graph_begin(void* stream);
int wpc_graph_end(void* stream, void** graph_exec);
int wpc_graph_launch(void* graph_exec, void* stream);
int wpc_graph_destroy(void* graph_exec);
/* A non-blocking stream of the caller's own, default priority (a framework's
 * pooled streams can alias; a capture needs a stream no other work uses). */
int wpc_stream_create(void** stream);
int wpc_stream_destroy(void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WPC_H */
