// Streamed CSR products: the C entry points (kernels in stream_impl.h, one
// translation unit per vector count in stream_nv<N>.hip).

#include "stream_impl.h"

using namespace wpc;

static int stream_dispatch(int nvec, const StreamArgs& a);
// A column form the sliced products accept: int32, 16-bit offsets, or stencil
// deltas (natural row order only).
static bool column_form_ok(int col16, const int* rows) {
    return col16 >= kColI32 && col16 <= kColStencil && (col16 != kColStencil || rows == nullptr);
}
static int sqlap_cg_dispatch(int nvec, const SqlapCgArgs& a);

extern "C" {

int wpc_csr_row_blocks(int n, const int* host_offsets, int* host_blocks, int capacity) {
    if (n < 0 || capacity < 2) return WPC_EARG;
    int count = 0;
    int r = 0;
    host_blocks[0] = 0;
    while (r < n) {
        const int start = r;
        const int e0 = host_offsets[r];
        int rows = 0;
        while (r < n && rows < kStreamRows) {
            const int len = host_offsets[r + 1] - host_offsets[r];
            const int used = host_offsets[r] - e0;
            if (rows > 0 && used + len > kStreamNnz) break;
            ++r;
            ++rows;
            if (len > kStreamNnz) break;  // a long row stands alone
        }
        (void)start;
        if (count + 2 > capacity) return WPC_ESCRATCH;
        host_blocks[++count] = r;
    }
    return count;
}

int wpc_csr_stream(int nblocks, const int* blocks, int n, int lanes, const int* offsets,
                   const int* columns, const void* values, int val_dtype, const void* s,
                   int s_dtype, const void* d, const void* u, const void* x, const void* y,
                   void* z, int vec_dtype, int nvec, double alpha, double beta, const void* w,
                   double* partials, double* out, void* stream) {
    int lg = -1;
    if (lanes > 0) {
        lg = 0;
        while ((1 << lg) < lanes) ++lg;
        if ((1 << lg) != lanes || lg > 6) return WPC_EARG;
    } else if (blocks == nullptr) {
        return WPC_EARG;
    }
    if (s != nullptr && s_dtype != WPC_F32 && s_dtype != vec_dtype) return WPC_EDTYPE;
    const StreamArgs a{nblocks, blocks, nullptr, nullptr, nullptr, false, n, lg, offsets, columns,
                       values, val_dtype, s,
                       s != nullptr && s_dtype == WPC_F32, d, u, x, y, z, vec_dtype, alpha, beta, w,
                       partials, out, as_stream(stream)};
    return stream_dispatch(nvec, a);
}

int wpc_sell_stream(int n, const int* slice_ptr, const int* slice_len, const int* slice_base,
                    const int* rows, const void* sell_columns, int col16,
                    const void* sell_values, int val_dtype,
                    const void* s, int s_dtype, const void* d, const void* u, const void* x,
                    const void* y, void* z, int vec_dtype, int nvec, double alpha, double beta,
                    const void* w, double* partials, double* out, void* stream) {
    if (n > 0 && (slice_ptr == nullptr || slice_len == nullptr)) return WPC_EARG;
    if (n > 0 && slice_base == nullptr) return WPC_EARG;
    if (!column_form_ok(col16, rows)) return WPC_EARG;
    if (s != nullptr && s_dtype != WPC_F32 && s_dtype != vec_dtype) return WPC_EDTYPE;
    const StreamArgs a{0, slice_ptr, slice_len, slice_base, rows, col16, n, -2, nullptr,
                       (const int*)sell_columns, sell_values,
                       val_dtype, s, s != nullptr && s_dtype == WPC_F32, d, u, x, y, z, vec_dtype,
                       alpha, beta, w, partials, out, as_stream(stream)};
    return stream_dispatch(nvec, a);
}

int wpc_sqlap_cg_partials(int n) {
    return n <= 0 ? 0 : (int)blocks_for(n, kSqlapCgRows);
}

int wpc_sqlap_cg_update(int n, const int* slice_ptr, const int* slice_len,
                        const int* slice_base, const int* rows, const void* sell_columns,
                        int col16, const void* sell_values, int val_dtype,
                        const void* inverse_mass, const void* weighted_mass, const void* p,
                        void* t, void* r, int vec_dtype, int nvec, double* part_energy,
                        double* part_rr, double* state, int passes, void* x, void* stream) {
    if (passes < 1 || passes > 7 || passes == 4 || passes == 5) return WPC_EARG;
    if (n <= 0) return WPC_OK;
    if (slice_ptr == nullptr || slice_len == nullptr || slice_base == nullptr ||
        sell_columns == nullptr || sell_values == nullptr || inverse_mass == nullptr ||
        p == nullptr || t == nullptr || r == nullptr || part_energy == nullptr ||
        part_rr == nullptr || state == nullptr)
        return WPC_EARG;
    if (p == t || p == r || t == r) return WPC_EARG;
    if (!column_form_ok(col16, rows)) return WPC_EARG;
    if ((passes & 4) && x == nullptr) return WPC_EARG;
    if (x != nullptr && (x == p || x == t || x == r)) return WPC_EARG;
    const SqlapCgArgs a{n, slice_ptr, slice_len, slice_base, rows, col16, sell_columns,
                        sell_values, val_dtype, inverse_mass, weighted_mass, p, t, r, vec_dtype,
                        part_energy, part_rr, state, passes, x, nullptr, 0, nullptr,
                        as_stream(stream)};
    return sqlap_cg_dispatch(nvec, a);
}

int wpc_sqlap_cg_direction_energy(int n, const int* slice_ptr, const int* slice_len,
                                  const int* slice_base, const int* rows,
                                  const void* sell_columns, int col16, const void* sell_values,
                                  int val_dtype, const void* inverse_mass,
                                  const void* weighted_mass, const void* z, void* p, void* t,
                                  int vec_dtype, int nvec, int nparts_rz, const double* part_rz,
                                  const double* part_rr, double* part_energy, double* state,
                                  void* stream) {
    if (n <= 0) return WPC_OK;
    if (slice_ptr == nullptr || slice_len == nullptr || slice_base == nullptr ||
        sell_columns == nullptr || sell_values == nullptr || inverse_mass == nullptr ||
        z == nullptr || p == nullptr || t == nullptr || part_energy == nullptr ||
        state == nullptr)
        return WPC_EARG;
    // part_rz NULL: BETA from the state (wpc_cg_beta); else both partial sets folded
    const bool from_state = part_rz == nullptr;
    if (!from_state && (part_rr == nullptr || nparts_rz <= 0)) return WPC_EARG;
    if (z == p || z == t || p == t) return WPC_EARG;
    if (!column_form_ok(col16, rows)) return WPC_EARG;
    const SqlapCgArgs a{n, slice_ptr, slice_len, slice_base, rows, col16, sell_columns,
                        sell_values, val_dtype, inverse_mass, weighted_mass, p, t, nullptr,
                        vec_dtype, part_energy, const_cast<double*>(part_rr), state,
                        1 | 8 | (from_state ? 4 : 0), nullptr, z, from_state ? 0 : nparts_rz,
                        part_rz, as_stream(stream)};
    return sqlap_cg_dispatch(nvec, a);
}

int wpc_sell_stream_grouped(int ntasks, const void* tasks, int total_blocks, int val_dtype,
                            int s_dtype, int has_d, int dot, int vec_dtype, double alpha,
                            void* stream) {
    if (ntasks <= 0) return WPC_OK;
    if (tasks == nullptr || total_blocks <= 0) return WPC_EARG;
    if (s_dtype != 0 && s_dtype != WPC_F32 && s_dtype != vec_dtype) return WPC_EDTYPE;
    const GroupedArgs a{ntasks, (const SellTask*)tasks, total_blocks, val_dtype, s_dtype != 0,
                        s_dtype == WPC_F32, has_d != 0, dot != 0, vec_dtype, alpha,
                        as_stream(stream)};
    return stream_grouped_entry(a);
}

}  // extern "C"

static int sqlap_cg_dispatch(int nvec, const SqlapCgArgs& a) {
    switch (nvec) {
        case 1: return sqlap_cg_entry_1(a);
        case 2: return sqlap_cg_entry_2(a);
        case 3: return sqlap_cg_entry_3(a);
        case 4: return sqlap_cg_entry_4(a);
        case 5: return sqlap_cg_entry_5(a);
        case 6: return sqlap_cg_entry_6(a);
        case 7: return sqlap_cg_entry_7(a);
        case 8: return sqlap_cg_entry_8(a);
        default: return WPC_EARG;
    }
}

static int stream_dispatch(int nvec, const StreamArgs& a) {
    switch (nvec) {
        case 1: return stream_entry_1(a);
        case 2: return stream_entry_2(a);
        case 3: return stream_entry_3(a);
        case 4: return stream_entry_4(a);
        case 5: return stream_entry_5(a);
        case 6: return stream_entry_6(a);
        case 7: return stream_entry_7(a);
        case 8: return stream_entry_8(a);
        default: return WPC_EARG;
    }
}
