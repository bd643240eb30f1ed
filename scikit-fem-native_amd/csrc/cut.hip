// Cut-cell (ragged quadrature) assembly for gfx950 (MI355X).
//
// Replaces CutLinearFormAssembler / CutBilinearFormAssembler
// (cpp/src/assembly/cut_form_assembler.cpp:67-239) and CutCrossAssembler
// (cpp/src/assembly/cut_cross_assembler.cpp:16-165).  A cut cell carries its
// own run of quadrature points [offsets[e], offsets[e+1]) with tabulated
// shape values and physical gradients per point, so there is no geometry to
// compute: the work per entry is a segmented sum over the cell's points.
//
// Execution model.  The rectangular plan of the cut cells' node
// connectivity (cells without points left out, as the reference's pattern
// leaves them out) gives every row node its incidences in cell order, the
// node-level pattern and, per incidence, the in-row offsets of the cell's
// column nodes.  One 64-lane wave owns G consecutive row nodes; each has a
// lane group of 64/G lanes and lane j of a group owns the cell's column
// nodes j, j + 64/G, ...  For each incidence the lane adds its entries (for
// value / gradient terms the cell's node matrix, formed beforehand one
// wave per cell on the matrix cores; for tensor terms and cells that repeat a node, its sum
// over the cell's points) into the wave's CSR span in LDS.  A
// lane's entries have distinct addresses (distinct column nodes; cells that
// repeat a node run with one lane per row), groups own disjoint spans, and
// every entry adds its incidences in cell order, so results are
// reproducible.  The span is then written to HBM once, coalesced.  The load
// vector uses the same incidences with one lane per (row node, component).
#include <hip/hip_runtime.h>
#include <algorithm>

#include "skfa_internal.h"
#include "skfa_plan.h"
#include "../../include/skfem_amd.h"

namespace skfa {
namespace {

struct CutArgs {
    const int64_t* inc_ptr;
    const int64_t* nptr;
    const uint8_t* records;
    int32_t rec_bytes, off_bytes;
    int64_t row_nodes;
    int32_t nr, nc, Cr, Cc;
    int32_t rows_per_wave, group;
    const int64_t* offsets;  // (cells + 1) point ranges of the plan's cells
    const double* weights;   // (P)
    const double* rshape;    // (P, nr)
    const double* rgrad;     // (P, nr, Dr)
    const double* cshape;    // (P, nc)
    const double* cgrad;     // (P, nc, Dc)
    int32_t Dr, Dc;          // spatial dimensions of the two gradient tables
    int32_t terms;           // SKFA_CUT_VALUE | SKFA_CUT_GRADIENT, or SKFA_CUT_TENSOR
    const double* value;     // (P) or NULL = 1
    const double* gradient;  // (P) or NULL = 1
    int32_t row_grad, col_grad;
    const double* tensor;    // (P, Cr, [Dr], Cc, [Dc])
    double* values;
};

__device__ inline int rec_offset(const uint8_t* rec, int off_bytes, int b) {
    return off_bytes == 1 ? int(rec[4 + b]) : int(reinterpret_cast<const uint16_t*>(rec + 4)[b]);
}

__global__ void __launch_bounds__(64) k_cut_rows(CutArgs A) {
    extern __shared__ double span[];
    const int lane = threadIdx.x;
    const int slot = lane / A.group;
    const int j0 = lane - slot * A.group;
    const int64_t r0 = int64_t(blockIdx.x) * A.rows_per_wave;
    const int64_t r1 = min(r0 + int64_t(A.rows_per_wave), A.row_nodes);
    const int64_t cc = int64_t(A.Cr) * A.Cc;
    const int64_t lo = A.nptr[r0];
    const int64_t span_len = (A.nptr[r1] - lo) * cc;
    for (int64_t k = lane; k < span_len; k += 64) span[k] = 0.;
    __syncthreads();
    const int64_t r = r0 + slot;
    if (r < r1) {
        const int64_t len = A.nptr[r + 1] - A.nptr[r];
        double* rs = span + (A.nptr[r] - lo) * cc;
        const bool tensor = A.terms == SKFA_CUT_TENSOR;
        const int Kr = A.row_grad ? A.Dr : 1, Kc = A.col_grad ? A.Dc : 1;
        for (int64_t k = A.inc_ptr[r]; k < A.inc_ptr[r + 1]; ++k) {
            const uint8_t* rec = A.records + k * A.rec_bytes;
            const int32_t packed = *reinterpret_cast<const int32_t*>(rec);
            const int64_t e = packed >> 4;
            const int a = packed & 15;
            const int64_t q0 = A.offsets[e], q1 = A.offsets[e + 1];
            for (int b = j0; b < A.nc; b += A.group) {
                const int o = rec_offset(rec, A.off_bytes, b);
                if (!tensor) {
                    // component-diagonal entries (a, c), (b, c) share one value
                    double acc = 0.;
                    for (int64_t q = q0; q < q1; ++q) {
                        double v = 0.;
                        if (A.terms & SKFA_CUT_VALUE)
                            v += (A.value ? A.value[q] : 1.) * A.rshape[q * A.nr + a] * A.cshape[q * A.nc + b];
                        if (A.terms & SKFA_CUT_GRADIENT) {
                            const double* ga = A.rgrad + (q * A.nr + a) * A.Dr;
                            const double* gb = A.cgrad + (q * A.nc + b) * A.Dc;
                            double dot = 0.;
                            for (int d = 0; d < A.Dr; ++d) dot += ga[d] * gb[d];
                            v += (A.gradient ? A.gradient[q] : 1.) * dot;
                        }
                        acc += v * A.weights[q];
                    }
                    for (int c = 0; c < A.Cr; ++c) rs[(int64_t(c) * len + o) * A.Cc + c] += acc;
                } else {
                    for (int cr = 0; cr < A.Cr; ++cr)
                        for (int ccol = 0; ccol < A.Cc; ++ccol) {
                            double acc = 0.;
                            for (int64_t q = q0; q < q1; ++q) {
                                const double* T = A.tensor + q * (int64_t(A.Cr) * Kr * A.Cc * Kc);
                                double v = 0.;
                                for (int i = 0; i < Kr; ++i) {
                                    const double ha = A.row_grad ? A.rgrad[(q * A.nr + a) * A.Dr + i]
                                                                 : A.rshape[q * A.nr + a];
                                    for (int jj = 0; jj < Kc; ++jj) {
                                        const double hb = A.col_grad ? A.cgrad[(q * A.nc + b) * A.Dc + jj]
                                                                     : A.cshape[q * A.nc + b];
                                        v += ha * T[((cr * Kr + i) * A.Cc + ccol) * Kc + jj] * hb;
                                    }
                                }
                                acc += A.weights[q] * v;
                            }
                            rs[(int64_t(cr) * len + o) * A.Cc + ccol] += acc;
                        }
                }
            }
        }
    }
    __syncthreads();
    for (int64_t k = lane; k < span_len; k += 64) A.values[lo * cc + k] = span[k];
}

// Value / gradient terms on cells without repeated nodes, in two passes.
// k_cut_cell_matrices: one wave per cell forms the cell's node matrix
//   M_e[a][b] = sum_q w_q (value_q phi_a(q) psi_b(q)
//                          + gradient_q grad phi_a(q) . grad psi_b(q))
// (the per-point entry of cut_form_assembler.cpp:182-190, summed per cell)
// as a product of a (n_row x 4Q) and a (4Q x n_col) matrix, one FP64
// 16x16x4 MFMA per point: lane l supplies slice k = l / 16 (0 the value
// term, 1..3 a gradient component) of row a = l % 16 of the left operand,
//   w_q value_q phi_a(q)  or  w_q gradient_q d_k phi_a(q),
// and of column b = l % 16 of the right operand, psi_b(q) or d_k psi_b(q),
// and receives entries M[l / 16 + 4r][l % 16].  Each lane loads one value
// of each operand per point, so a cell's point rows pass through the cache
// once instead of once per entry.  NP points issue their loads together.
// k_cut_rows_cells: a lane group of `group` lanes per row node, lane b
// owning column node b; each incidence adds M_e[a][b] to the row's span
// entry (diagonal component pairs).  kCutBatch incidences of a row issue
// their record loads together and then their matrix loads together, so a
// row pays two memory round trips per batch instead of one per point.  The
// adds of a row run in incidence (cell) order, one lane per address.
typedef double cut_v4d __attribute__((ext_vector_type(4)));
constexpr int kCutBatch = 8;

__device__ inline double read_lane(double x, int l) {
    const int64_t bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(int(bits), l);
    const int hi = __builtin_amdgcn_readlane(int(bits >> 32), l);
    return __longlong_as_double((int64_t(hi) << 32) | uint32_t(lo));
}

// SAME: row and column tables are the same arrays (square forms), so the
// right operand is the left one before its point factor.
template <bool SAME, int NP>
__global__ void __launch_bounds__(256) k_cut_cell_matrices(CutArgs A, double* cell_mat, int64_t cells) {
    const int lane = threadIdx.x & 63;
    const int64_t e = int64_t(blockIdx.x) * 4 + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    if (e >= cells) return;
    const int i = lane & 15, k = lane >> 4;
    const bool has_value = (A.terms & SKFA_CUT_VALUE) != 0, has_grad = (A.terms & SKFA_CUT_GRADIENT) != 0;
    const bool a_on = i < A.nr && (k == 0 ? has_value : (has_grad && k <= A.Dr));
    const bool b_on = i < A.nc && (k == 0 ? has_value : (has_grad && k <= A.Dc));
    const double* ah = k == 0 ? A.rshape + i : A.rgrad + i * A.Dr + (k - 1);
    const int64_t as = k == 0 ? A.nr : int64_t(A.nr) * A.Dr;
    const double* bh = k == 0 ? A.cshape + i : A.cgrad + i * A.Dc + (k - 1);
    const int64_t bs = k == 0 ? A.nc : int64_t(A.nc) * A.Dc;
    const int64_t q0 = A.offsets[e], q1 = A.offsets[e + 1];
    cut_v4d acc = {0., 0., 0., 0.};
    for (int64_t qc = q0; qc < q1; qc += 64) {
        // point qc + lane's factors w value and w gradient, read by every
        // lane below through readlane
        const int64_t ql = qc + lane;
        double fv = 0., fg = 0.;
        if (ql < q1) {
            const double w = A.weights[ql];
            fv = w * (A.value ? A.value[ql] : 1.);
            fg = w * (A.gradient ? A.gradient[ql] : 1.);
        }
        const int64_t qe = min(qc + 64, q1);
        for (int64_t qb = qc; qb < qe; qb += NP) {
            double hv[NP], bv[SAME ? 1 : NP];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const int64_t q = qb + j;
                hv[j] = 0.;
                if (!SAME) bv[j] = 0.;
                if (q < qe) {
                    if (a_on) hv[j] = ah[q * as];
                    if (!SAME && b_on) bv[j] = bh[q * bs];
                }
            }
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const int src = int(qb - qc) + j;   // uniform; beyond qe the operands are 0
                const double f = k == 0 ? read_lane(fv, src & 63) : read_lane(fg, src & 63);
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(f * hv[j], SAME ? hv[j] : bv[SAME ? 0 : j], acc, 0, 0, 0);
            }
        }
    }
    double* out = cell_mat + e * (int64_t(A.nr) * A.nc);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int a = k + 4 * r;
        if (a < A.nr && i < A.nc) out[a * A.nc + i] = acc[r];
    }
}

__global__ void __launch_bounds__(64) k_cut_rows_cells(CutArgs A, const double* cell_mat) {
    extern __shared__ double span[];
    const int lane = threadIdx.x;
    const int slot = lane / A.group;
    const int b = lane - slot * A.group;
    const int64_t r0 = int64_t(blockIdx.x) * A.rows_per_wave;
    const int64_t r1 = min(r0 + int64_t(A.rows_per_wave), A.row_nodes);
    const int64_t cc = int64_t(A.Cr) * A.Cc;
    const int64_t lo = A.nptr[r0];
    const int64_t span_len = (A.nptr[r1] - lo) * cc;
    for (int64_t k = lane; k < span_len; k += 64) span[k] = 0.;
    __syncthreads();
    const int64_t r = r0 + slot;
    if (r < r1 && b < A.nc) {
        const int64_t len = A.nptr[r + 1] - A.nptr[r];
        double* rs = span + (A.nptr[r] - lo) * cc;
        const int64_t nn = int64_t(A.nr) * A.nc;
        const int64_t kb = A.inc_ptr[r], ke = A.inc_ptr[r + 1];
        for (int64_t k = kb; k < ke; k += kCutBatch) {
            int32_t pk[kCutBatch];
            int o[kCutBatch];
            double m[kCutBatch];
#pragma unroll
            for (int j = 0; j < kCutBatch; ++j) {
                pk[j] = 0;
                o[j] = 0;
                if (k + j < ke) {
                    const uint8_t* rec = A.records + (k + j) * A.rec_bytes;
                    pk[j] = *reinterpret_cast<const int32_t*>(rec);
                    o[j] = rec_offset(rec, A.off_bytes, b);
                }
            }
#pragma unroll
            for (int j = 0; j < kCutBatch; ++j)
                m[j] = k + j < ke ? cell_mat[int64_t(pk[j] >> 4) * nn + (pk[j] & 15) * A.nc + b] : 0.;
#pragma unroll
            for (int j = 0; j < kCutBatch; ++j)
                if (k + j < ke)
                    for (int c = 0; c < A.Cr; ++c) rs[(int64_t(c) * len + o[j]) * A.Cc + c] += m[j];
        }
    }
    __syncthreads();
    for (int64_t k = lane; k < span_len; k += 64) A.values[lo * cc + k] = span[k];
}

struct CutLinearArgs {
    const int64_t* inc_ptr;
    const uint8_t* records;
    int32_t rec_bytes;
    int64_t row_nodes;
    int32_t n, C, D;
    const int64_t* offsets;
    const double* weights;
    const double* shape;     // (P, n)
    const double* grad;      // (P, n, D)
    const double* value;     // (C) or (P, C), NULL = no value term
    int32_t value_per_point;
    const double* gradient;  // (C, D) or (P, C, D), NULL = no gradient term
    int32_t gradient_per_point;
    double* out;             // (row_nodes * C)
};

// Load vector in two passes, as the bilinear value / gradient terms.
// k_cut_cell_vectors: one thread per (cell, node a, component c) sums
//   b_e[a][c] = sum_q w_q (phi_a(q) f_c(q) + grad phi_a(q) . G_c(q))
// over the cell's points in order (the per-point contribution of
// cut_form_assembler.cpp:124-137, summed per cell); threads of one cell
// read different shape rows, so the point tables pass through once.
// k_cut_linear_rows: one lane per DOF row (node * C + c) adds b_e[a][c]
// over the node's incidences in cell order, kCutBatch incidences' records
// and cell values loaded together.
__global__ void __launch_bounds__(256) k_cut_cell_vectors(CutLinearArgs A, double* cell_vec, int64_t total) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int nC = A.n * A.C;
    const int64_t e = t / nC;
    const int r = int(t - e * nC);
    const int a = r / A.C, c = r - a * A.C;
    double s = 0.;
#pragma unroll 4
    for (int64_t q = A.offsets[e]; q < A.offsets[e + 1]; ++q) {
        double con = 0.;
        if (A.value) con += A.shape[q * A.n + a] * A.value[A.value_per_point ? q * A.C + c : c];
        if (A.gradient) {
            const double* g = A.grad + (q * A.n + a) * A.D;
            const double* G = A.gradient + (A.gradient_per_point ? (q * A.C + c) * A.D : int64_t(c) * A.D);
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (d < A.D) con += g[d] * G[d];
        }
        s += A.weights[q] * con;
    }
    cell_vec[t] = s;
}

__global__ void __launch_bounds__(256) k_cut_linear_rows(CutLinearArgs A, const double* cell_vec) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= A.row_nodes * A.C) return;
    const int64_t node = t / A.C;
    const int c = int(t - node * A.C);
    const int64_t kb = A.inc_ptr[node], ke = A.inc_ptr[node + 1];
    double acc = 0.;
    for (int64_t k = kb; k < ke; k += kCutBatch) {
        int32_t pk[kCutBatch];
        double v[kCutBatch];
#pragma unroll
        for (int j = 0; j < kCutBatch; ++j)
            pk[j] = k + j < ke ? *reinterpret_cast<const int32_t*>(A.records + (k + j) * A.rec_bytes) : 0;
#pragma unroll
        for (int j = 0; j < kCutBatch; ++j)
            v[j] = k + j < ke ? cell_vec[(int64_t(pk[j] >> 4) * A.n + (pk[j] & 15)) * A.C + c] : 0.;
#pragma unroll
        for (int j = 0; j < kCutBatch; ++j)
            if (k + j < ke) acc += v[j];
    }
    A.out[t] = acc;
}

}  // namespace
}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_assemble_cut_bilinear(const skfa_cross_plan* P, const int64_t* offsets, const double* weights,
                               const double* row_shape, const double* row_grad, int32_t row_dim,
                               const double* col_shape, const double* col_grad, int32_t col_dim, int32_t terms,
                               const double* value, const double* gradient, int32_t row_kind, int32_t col_kind,
                               const double* tensor, int32_t one_lane_per_row, double* values, void* stream) {
    SKFA_REQUIRE(P != nullptr, "plan is null");
    SKFA_REQUIRE(values != nullptr || P->nnz() == 0, "values pointer is null");
    SKFA_REQUIRE(row_dim >= 1 && row_dim <= 3 && col_dim >= 1 && col_dim <= 3, "spatial dimension must be 1, 2 or 3");
    SKFA_REQUIRE(row_kind == SKFA_CROSS_VALUE || row_kind == SKFA_CROSS_GRADIENT, "row kind must be value or gradient");
    SKFA_REQUIRE(col_kind == SKFA_CROSS_VALUE || col_kind == SKFA_CROSS_GRADIENT,
                 "column kind must be value or gradient");
    if (terms == SKFA_CUT_TENSOR) {
        SKFA_REQUIRE(tensor != nullptr, "tensor coefficient is null");
    } else {
        SKFA_REQUIRE(terms >= 1 && terms <= (SKFA_CUT_VALUE | SKFA_CUT_GRADIENT), "at least one term is required");
        SKFA_REQUIRE(P->C_row == P->C_col, "scalar cut coefficient requires equal component counts");
        SKFA_REQUIRE(!(terms & SKFA_CUT_GRADIENT) || row_dim == col_dim,
                     "scalar cut gradient coefficient requires equal dimensions");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->nnz() == 0 || P->row_nodes == 0) return kOk;
    SKFA_REQUIRE(offsets != nullptr && weights != nullptr && row_shape != nullptr && col_shape != nullptr,
                 "cut point data pointer is null");
    SKFA_REQUIRE(row_grad != nullptr && col_grad != nullptr, "cut gradient table pointer is null");
    const int64_t row_bytes = int64_t(P->max_row_len) * P->C_row * P->C_col * 8;
    if (row_bytes > kMaxWaveLds) {
        set_error("the CSR span of one row node exceeds 64 KiB of LDS");
        return kCapacity;
    }
    int group = 1;
    if (!one_lane_per_row)
        while (group < P->n_col && group < 64) group <<= 1;
    int rows = 64 / group;
    while (rows > 1 && rows * row_bytes > kMaxWaveLds) rows >>= 1;
    CutArgs A{};
    A.inc_ptr = P->inc_ptr.p;
    A.nptr = P->indptr.p;
    A.records = P->records.p;
    A.rec_bytes = P->rec_bytes;
    A.off_bytes = P->off_bytes;
    A.row_nodes = P->row_nodes;
    A.nr = P->n_row;
    A.nc = P->n_col;
    A.Cr = P->C_row;
    A.Cc = P->C_col;
    A.rows_per_wave = rows;
    A.group = 64 / rows;
    A.offsets = offsets;
    A.weights = weights;
    A.rshape = row_shape;
    A.rgrad = row_grad;
    A.cshape = col_shape;
    A.cgrad = col_grad;
    A.Dr = row_dim;
    A.Dc = col_dim;
    A.terms = terms;
    A.value = value;
    A.gradient = gradient;
    A.row_grad = row_kind == SKFA_CROSS_GRADIENT;
    A.col_grad = col_kind == SKFA_CROSS_GRADIENT;
    A.tensor = tensor;
    A.values = values;
    const int64_t waves = (P->row_nodes + rows - 1) / rows;
    SKFA_REQUIRE(waves <= int64_t(INT32_MAX), "too many row nodes");
    // value / gradient terms on cells without repeated nodes: per-cell node
    // matrices, then the row gather (a lane per column node)
    if (terms != SKFA_CUT_TENSOR && !one_lane_per_row) {
        const int64_t total = P->E * int64_t(P->n_row) * P->n_col;
        const int64_t blocks = (P->E + 3) / 4;   // a wave per cell
        SKFA_REQUIRE(blocks <= int64_t(INT32_MAX), "too many cut cells");
        if (P->cell_buffer.n < size_t(total)) {
            const int rc = P->cell_buffer.alloc(size_t(total));
            if (rc != kOk) return rc;
        }
        const bool same = row_shape == col_shape && row_grad == col_grad && P->n_row == P->n_col && row_dim == col_dim;
        if (same)
            k_cut_cell_matrices<true, 8><<<dim3(unsigned(blocks)), dim3(256), 0, s>>>(A, P->cell_buffer.p, P->E);
        else
            k_cut_cell_matrices<false, 8><<<dim3(unsigned(blocks)), dim3(256), 0, s>>>(A, P->cell_buffer.p, P->E);
        SKFA_LAUNCH_CHECK();
        k_cut_rows_cells<<<dim3(unsigned(waves)), dim3(64), size_t(rows * row_bytes), s>>>(A, P->cell_buffer.p);
        SKFA_LAUNCH_CHECK();
        return kOk;
    }
    k_cut_rows<<<dim3(unsigned(waves)), dim3(64), size_t(rows * row_bytes), s>>>(A);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_assemble_cut_linear(const skfa_cross_plan* P, const int64_t* offsets, const double* weights,
                             const double* shape, const double* grad, int32_t dim, const double* value,
                             int32_t value_per_point, const double* gradient, int32_t gradient_per_point,
                             double* out, void* stream) {
    SKFA_REQUIRE(P != nullptr, "plan is null");
    SKFA_REQUIRE(value != nullptr || gradient != nullptr, "at least one coefficient is required");
    SKFA_REQUIRE(dim >= 1 && dim <= 3, "spatial dimension must be 1, 2 or 3");
    const int64_t total = P->row_nodes * P->C_row;
    if (total == 0) return kOk;
    SKFA_REQUIRE(out != nullptr, "output pointer is null");
    SKFA_REQUIRE(P->E == 0 || (offsets != nullptr && weights != nullptr && shape != nullptr && grad != nullptr),
                 "cut point data pointer is null");
    CutLinearArgs A{};
    A.inc_ptr = P->inc_ptr.p;
    A.records = P->records.p;
    A.rec_bytes = P->rec_bytes;
    A.row_nodes = P->row_nodes;
    A.n = P->n_row;
    A.C = P->C_row;
    A.D = dim;
    A.offsets = offsets;
    A.weights = weights;
    A.shape = shape;
    A.grad = grad;
    A.value = value;
    A.value_per_point = value_per_point;
    A.gradient = gradient;
    A.gradient_per_point = gradient_per_point;
    A.out = out;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t blocks = (total + 255) / 256;
    SKFA_REQUIRE(blocks <= int64_t(INT32_MAX), "too many rows");
    const int64_t cells = P->E * int64_t(P->n_row) * P->C_row;
    const int64_t cblocks = (cells + 255) / 256;
    SKFA_REQUIRE(cblocks <= int64_t(INT32_MAX), "too many cut cells");
    if (cells > 0) {
        if (P->cell_buffer.n < size_t(cells)) {
            const int rc = P->cell_buffer.alloc(size_t(cells));
            if (rc != kOk) return rc;
        }
        k_cut_cell_vectors<<<dim3(unsigned(cblocks)), dim3(256), 0, s>>>(A, P->cell_buffer.p, cells);
        SKFA_LAUNCH_CHECK();
    }
    k_cut_linear_rows<<<dim3(unsigned(blocks)), dim3(256), 0, s>>>(A, P->cell_buffer.p);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

}  // extern "C"
