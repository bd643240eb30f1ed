// Rectangular (cross-basis) element assembly for gfx950 (MI355X).
//
// Replaces CrossBilinearAssembler::assemble for two nodal fields on one
// tetrahedral mesh (cpp/src/assembly/cross_bilinear_assembler.cpp:55-107,
// 207-247; contractions cpp/include/native_fem/cross_contraction.hpp:12-97):
// Taylor-Hood divergence blocks, value and gradient couplings between P1 and
// P2 fields, and caller-supplied contraction tensors.
//
// Execution model.  One 64-lane wave owns G consecutive row (test) nodes;
// each node has a lane group of 64/G lanes, and lane j of a group owns the
// local entries (row component r, column node b, column component c) with
// index j, j + 64/G, ...  The CSR values of the wave's rows are one
// contiguous span (row DOF node*Cr + r holds its node's column nodes times
// Cc, node-major), accumulated in LDS.  A group walks its node's incidences
// in element order; per incidence every lane rebuilds the element's
// geometry (affine Jacobian once, or the isoparametric one per quadrature
// point), integrates its entries over the quadrature points and adds them
// at the record's in-row node offsets.  Entries of one incidence hit
// distinct addresses and the groups own disjoint spans, so the sums follow
// a fixed order and results are reproducible.  The span is then written to
// HBM once, coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "skfa_internal.h"
#include "skfa_plan.h"
#include "../../include/skfem_amd.h"

namespace skfa {
namespace {

struct CrossArgs {
    const int64_t* inc_ptr;   // row-node incidences (records in the same order)
    const int64_t* nptr;      // node-level rectangular indptr
    const uint8_t* records;
    int32_t rec_bytes, off_bytes;
    int64_t row_nodes;
    int32_t nr, nc, Cr, Cc;
    int32_t rows_per_wave, group;  // G and 64 / G
    const double4* xyz;
    const int32_t* vconn;
    const int32_t* gconn;
    int32_t gn;
    const double* tables;  // w | rphi | rdphi | cphi | cdphi | gdphi
    int32_t Q;
    int32_t row_grad, col_grad, mode;
    const double* coef;
    int64_t cse, csq;
    double* values;
};

struct Jac {
    double adj[9];  // adjugate (J^{-1} = adj / det), row-major
    double det;
};

__device__ inline void jac_affine(const double4* __restrict__ xyz, const int32_t* __restrict__ c4, Jac& g) {
    const double4 v0 = xyz[c4[0]], v1 = xyz[c4[1]], v2 = xyz[c4[2]], v3 = xyz[c4[3]];
    const double J[9] = {v1.x - v0.x, v2.x - v0.x, v3.x - v0.x, v1.y - v0.y, v2.y - v0.y,
                         v3.y - v0.y, v1.z - v0.z, v2.z - v0.z, v3.z - v0.z};
    g.adj[0] = J[4] * J[8] - J[5] * J[7];
    g.adj[1] = J[2] * J[7] - J[1] * J[8];
    g.adj[2] = J[1] * J[5] - J[2] * J[4];
    g.adj[3] = J[5] * J[6] - J[3] * J[8];
    g.adj[4] = J[0] * J[8] - J[2] * J[6];
    g.adj[5] = J[2] * J[3] - J[0] * J[5];
    g.adj[6] = J[3] * J[7] - J[4] * J[6];
    g.adj[7] = J[1] * J[6] - J[0] * J[7];
    g.adj[8] = J[0] * J[4] - J[1] * J[3];
    g.det = J[0] * g.adj[0] + J[1] * g.adj[3] + J[2] * g.adj[6];
}

// J_ij = sum_a x_{a,i} dN_a/dxi_j (basis_geometry.cpp:112-129)
__device__ inline void jac_iso(const double4* __restrict__ xyz, const int32_t* __restrict__ cg, int gn,
                               const double* __restrict__ gd, Jac& g) {
    double J[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
    for (int a = 0; a < gn; ++a) {
        const double4 x = xyz[cg[a]];
        const double g0 = gd[a * 3 + 0], g1 = gd[a * 3 + 1], g2 = gd[a * 3 + 2];
        J[0] += x.x * g0; J[1] += x.x * g1; J[2] += x.x * g2;
        J[3] += x.y * g0; J[4] += x.y * g1; J[5] += x.y * g2;
        J[6] += x.z * g0; J[7] += x.z * g1; J[8] += x.z * g2;
    }
    g.adj[0] = J[4] * J[8] - J[5] * J[7];
    g.adj[1] = J[2] * J[7] - J[1] * J[8];
    g.adj[2] = J[1] * J[5] - J[2] * J[4];
    g.adj[3] = J[5] * J[6] - J[3] * J[8];
    g.adj[4] = J[0] * J[8] - J[2] * J[6];
    g.adj[5] = J[2] * J[3] - J[0] * J[5];
    g.adj[6] = J[3] * J[7] - J[4] * J[6];
    g.adj[7] = J[1] * J[6] - J[0] * J[7];
    g.adj[8] = J[0] * J[4] - J[1] * J[3];
    g.det = J[0] * g.adj[0] + J[1] * g.adj[3] + J[2] * g.adj[6];
}

// Physical gradient dphi/dx_i = sum_k dphi/dxi_k (J^{-1})_{ki}
// (basis_geometry.cpp:161-173).
__device__ inline void phys_grad(const double* __restrict__ rg, const Jac& g, double inv, double* out) {
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = (rg[0] * g.adj[i] + rg[1] * g.adj[3 + i] + rg[2] * g.adj[6 + i]) * inv;
}

__global__ void __launch_bounds__(64) k_cross_rows(CrossArgs A) {
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
        const int ncc = A.nc * A.Cc;
        const int nent = A.Cr * ncc;
        const int Q = A.Q;
        const double* W = A.tables;
        const double* rphi = W + Q;
        const double* rdphi = rphi + Q * A.nr;
        const double* cphi = rdphi + Q * A.nr * 3;
        const double* cdphi = cphi + Q * A.nc;
        const double* gd = cdphi + Q * A.nc * 3;
        const int Dr = A.row_grad ? 3 : 1, Dc = A.col_grad ? 3 : 1;
        for (int64_t k = A.inc_ptr[r]; k < A.inc_ptr[r + 1]; ++k) {
            const uint8_t* rec = A.records + k * A.rec_bytes;
            const int32_t packed = *reinterpret_cast<const int32_t*>(rec);
            const int64_t e = packed >> 4;
            const int a = packed & 15;
            Jac G;
            if (!A.gconn) jac_affine(A.xyz, A.vconn + e * 4, G);
            for (int j = j0; j < nent; j += A.group) {
                const int cr = j / ncc;
                const int rem = j - cr * ncc;
                const int b = rem / A.Cc;
                const int ccp = rem - b * A.Cc;
                const int o = A.off_bytes == 1 ? int(rec[4 + b]) : int(reinterpret_cast<const uint16_t*>(rec + 4)[b]);
                // the scalar mode couples equal components only; the entry
                // stays in the pattern with value zero
                if (A.mode == SKFA_CROSS_SCALAR && cr != ccp) continue;
                double acc = 0.;
                for (int q = 0; q < Q; ++q) {
                    if (A.gconn) jac_iso(A.xyz, A.gconn + e * A.gn, A.gn, gd + q * A.gn * 3, G);
                    const double inv = 1. / G.det;
                    const double dx = fabs(G.det) * W[q];
                    double ra[3], cb[3];
                    if (A.row_grad) {
                        phys_grad(rdphi + (q * A.nr + a) * 3, G, inv, ra);
                    } else {
                        ra[0] = rphi[q * A.nr + a];
                        ra[1] = ra[2] = 0.;
                    }
                    if (A.col_grad) {
                        phys_grad(cdphi + (q * A.nc + b) * 3, G, inv, cb);
                    } else {
                        cb[0] = cphi[q * A.nc + b];
                        cb[1] = cb[2] = 0.;
                    }
                    double v;
                    if (A.mode == SKFA_CROSS_TENSOR) {
                        const double* T = A.coef + e * A.cse + q * A.csq;
                        v = 0.;
                        for (int i = 0; i < Dr; ++i)
                            for (int jj = 0; jj < Dc; ++jj)
                                v += ra[i] * T[((cr * Dr + i) * A.Cc + ccp) * Dc + jj] * cb[jj];
                    } else {
                        const double c = A.coef ? A.coef[e * A.cse + q * A.csq] : 1.;
                        double contract;
                        if (A.mode == SKFA_CROSS_ROW_DIVERGENCE)
                            contract = ra[cr] * cb[0];
                        else if (A.mode == SKFA_CROSS_COLUMN_DIVERGENCE)
                            contract = ra[0] * cb[ccp];
                        else if (A.row_grad)
                            contract = ra[0] * cb[0] + ra[1] * cb[1] + ra[2] * cb[2];
                        else
                            contract = ra[0] * cb[0];
                        v = c * contract;
                    }
                    acc += dx * v;
                }
                rs[(int64_t(cr) * len + o) * A.Cc + ccp] += acc;
            }
        }
    }
    __syncthreads();
    for (int64_t k = lane; k < span_len; k += 64) A.values[lo * cc + k] = span[k];
}

// Fast path: affine geometry, element-constant scalar coefficient, at most
// 32 entries per element row.  Lane (i, p) of a wave owns row node r0 + i
// and takes that node's incidences p, p + S, ... (S lanes per row), so 64
// incidences are in flight per wave.  An element row is evaluated from the
// reduced reference table rt[a][b][k][l] = sum_q w_q h_ak(q) h_bl(q) (h = phi
// or dphi/dxi) and the element's J^{-1}, |det J| and coefficient, kept in
// registers, and added to the span by the S parts in turn (one writer per
// address per ds_add_f64): every entry sums its incidences in element order,
// as in the general kernel.  Loads are software-pipelined: a lane's records
// run two steps ahead and its elements' corner ids one step ahead, so a
// step waits only for the four vertex coordinates.
// Gather record words (8, 16 or 32 bytes: 4 + NC offset bytes, rounded up).
struct RecWords {
    uint32_t w[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
};

__device__ inline RecWords load_rec(const uint8_t* __restrict__ records, int64_t k, int rec_bytes) {
    RecWords r;
    const uint8_t* base = records + k * rec_bytes;
    if (rec_bytes == 8) {
        const uint2 v = *reinterpret_cast<const uint2*>(base);
        r.w[0] = v.x;
        r.w[1] = v.y;
    } else {
        const uint4* src = reinterpret_cast<const uint4*>(base);
        const uint4 v0 = src[0];
        r.w[0] = v0.x; r.w[1] = v0.y; r.w[2] = v0.z; r.w[3] = v0.w;
        if (rec_bytes > 16) {
            const uint4 v1 = src[1];
            r.w[4] = v1.x; r.w[5] = v1.y; r.w[6] = v1.z; r.w[7] = v1.w;
        }
    }
    return r;
}

__device__ inline void jac_affine4(const double4* __restrict__ xyz, int4 c, Jac& g) {
    const double4 v0 = xyz[c.x], v1 = xyz[c.y], v2 = xyz[c.z], v3 = xyz[c.w];
    const double J[9] = {v1.x - v0.x, v2.x - v0.x, v3.x - v0.x, v1.y - v0.y, v2.y - v0.y,
                         v3.y - v0.y, v1.z - v0.z, v2.z - v0.z, v3.z - v0.z};
    g.adj[0] = J[4] * J[8] - J[5] * J[7];
    g.adj[1] = J[2] * J[7] - J[1] * J[8];
    g.adj[2] = J[1] * J[5] - J[2] * J[4];
    g.adj[3] = J[5] * J[6] - J[3] * J[8];
    g.adj[4] = J[0] * J[8] - J[2] * J[6];
    g.adj[5] = J[2] * J[3] - J[0] * J[5];
    g.adj[6] = J[3] * J[7] - J[4] * J[6];
    g.adj[7] = J[1] * J[6] - J[0] * J[7];
    g.adj[8] = J[0] * J[4] - J[1] * J[3];
    g.det = J[0] * g.adj[0] + J[1] * g.adj[3] + J[2] * g.adj[6];
}

// LDS span target of one fast-path wave: more rows per wave means fewer
// turns per row; measured on the 64^3 Taylor-Hood divergence block: 8 KiB
// 0.69 ms, 16 KiB 0.39 ms, 32 KiB 0.375 ms
constexpr int64_t kCrossLdsTarget = 32 * 1024;

struct CrossFastArgs {
    const int64_t* inc_ptr;
    const int64_t* nptr;
    const uint8_t* records;
    int32_t rec_bytes, off_bytes;
    int64_t row_nodes;
    int32_t S, R;
    const double4* xyz;
    const int32_t* vconn;
    const double* rtab;
    int32_t tab_len;
    int32_t lds_target;
    int32_t mode, row_grad, col_grad;
    const double* coef;
    int64_t cse;
    double* values;
};

template <int NR, int NC, int CR, int CC>
__global__ void __launch_bounds__(64) k_cross_fast(CrossFastArgs A) {
    constexpr int NENT = CR * NC * CC;
    static_assert(NENT <= 32, "fast cross rows hold at most 32 entries");
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int64_t r0 = int64_t(blockIdx.x) * A.R;
    const int64_t r1 = min(r0 + int64_t(A.R), A.row_nodes);
    constexpr int64_t cc = int64_t(CR) * CC;
    const int64_t lo = A.nptr[r0];
    const int64_t span_len = (A.nptr[r1] - lo) * cc;
    double* span = lds;
    double* tab = lds + span_len;
    for (int64_t k = lane; k < span_len; k += 64) span[k] = 0.;
    for (int k = lane; k < A.tab_len; k += 64) tab[k] = A.rtab[k];
    __syncthreads();
    const int i = lane / A.S;
    const int p = lane - i * A.S;
    const int64_t r = r0 + i;
    const bool active = r < r1;
    int64_t k0 = 0, k1 = 0, len = 0;
    double* rs = span;
    if (active) {
        k0 = A.inc_ptr[r];
        k1 = A.inc_ptr[r + 1];
        len = A.nptr[r + 1] - A.nptr[r];
        rs = span + (A.nptr[r] - lo) * cc;
    }
    const int DR = A.row_grad ? 3 : 1, DC = A.col_grad ? 3 : 1;
    // lanes of the wave run the same number of steps (uniform loop bound)
    int64_t steps = active ? (k1 - k0 + A.S - 1) / A.S : 0;
    for (int off = 32; off > 0; off >>= 1) steps = max(steps, (int64_t)__shfl_xor(steps, off, 64));
    // records of the next step are loaded while this step computes
    RecWords cw, nw;
    // records run two steps ahead, element corners one step ahead
    RecWords nnw;
    int4 cv = make_int4(0, 0, 0, 0), nv = make_int4(0, 0, 0, 0);
    int64_t kcur = k0 + p;
    bool hcur = active && kcur < k1;
    bool hnext = active && kcur + A.S < k1;
    if (hcur) cw = load_rec(A.records, kcur, A.rec_bytes);
    if (hnext) nw = load_rec(A.records, kcur + A.S, A.rec_bytes);
    if (hcur) cv = *reinterpret_cast<const int4*>(A.vconn + (int64_t(int32_t(cw.w[0])) >> 4) * 4);
    for (int64_t st = 0; st < steps; ++st) {
        const int64_t knext = kcur + A.S;
        const bool hnn = active && knext + A.S < k1;
        if (hnn) nnw = load_rec(A.records, knext + A.S, A.rec_bytes);
        if (hnext) nv = *reinterpret_cast<const int4*>(A.vconn + (int64_t(int32_t(nw.w[0])) >> 4) * 4);
        const bool have = hcur;
        double acc[NENT];
        int off_b[NC];
#pragma unroll
        for (int j = 0; j < NENT; ++j) acc[j] = 0.;
#pragma unroll
        for (int b = 0; b < NC; ++b) off_b[b] = 0;
        if (have) {
            const int32_t packed = int32_t(cw.w[0]);
            const int64_t e = packed >> 4;
            const int a = packed & 15;
            if (A.off_bytes == 1) {
#pragma unroll
                for (int b = 0; b < NC; ++b) off_b[b] = int((cw.w[1 + b / 4] >> (8 * (b % 4))) & 0xffu);
            } else {
#pragma unroll
                for (int b = 0; b < NC; ++b) off_b[b] = int((cw.w[1 + b / 2] >> (16 * (b % 2))) & 0xffffu);
            }
            Jac G;
            jac_affine4(A.xyz, cv, G);
            const double inv = 1. / G.det;
            const double s = fabs(G.det) * (A.coef ? A.coef[e * A.cse] : 1.);
            double Ji[9];  // J^{-1}
#pragma unroll
            for (int t = 0; t < 9; ++t) Ji[t] = G.adj[t] * inv;
            const double* ta = tab + a * NC * DR * DC;
            if (A.mode == SKFA_CROSS_SCALAR && !A.row_grad) {
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const double v = s * ta[b];
#pragma unroll
                    for (int c = 0; c < CR; ++c) acc[(c * NC + b) * CC + c] = v;
                }
            } else if (A.mode == SKFA_CROSS_SCALAR) {
                double Km[9];  // s J^{-1} J^{-T}
#pragma unroll
                for (int kk = 0; kk < 3; ++kk)
#pragma unroll
                    for (int ll = 0; ll < 3; ++ll)
                        Km[kk * 3 + ll] = s * (Ji[kk * 3] * Ji[ll * 3] + Ji[kk * 3 + 1] * Ji[ll * 3 + 1] +
                                               Ji[kk * 3 + 2] * Ji[ll * 3 + 2]);
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const double* t = ta + b * 9;
                    double v = 0.;
#pragma unroll
                    for (int kl = 0; kl < 9; ++kl) v += t[kl] * Km[kl];
#pragma unroll
                    for (int c = 0; c < CR; ++c) acc[(c * NC + b) * CC + c] = v;
                }
            } else if (A.mode == SKFA_CROSS_ROW_DIVERGENCE) {
                // entry (r, b): s sum_k rt[a][b][k] J^{-1}_{k r}
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const double* t = ta + b * 3;
#pragma unroll
                    for (int c = 0; c < CR; ++c)
                        acc[(c * NC + b) * CC] = s * (t[0] * Ji[c] + t[1] * Ji[3 + c] + t[2] * Ji[6 + c]);
                }
            } else {
                // column divergence, entry (b, c): s sum_l rt[a][b][l] J^{-1}_{l c}
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    const double* t = ta + b * 3;
#pragma unroll
                    for (int c = 0; c < CC; ++c)
                        acc[b * CC + c] = s * (t[0] * Ji[c] + t[1] * Ji[3 + c] + t[2] * Ji[6 + c]);
                }
            }
        }
        // the S parts of a row add in turn
        for (int turn = 0; turn < A.S; ++turn) {
            if (have && p == turn) {
#pragma unroll
                for (int c = 0; c < CR; ++c)
#pragma unroll
                    for (int b = 0; b < NC; ++b)
#pragma unroll
                        for (int d = 0; d < CC; ++d)
                            (void)atomicAdd(&rs[(int64_t(c) * len + off_b[b]) * CC + d], acc[(c * NC + b) * CC + d]);
            }
        }
        cw = nw;
        nw = nnw;
        cv = nv;
        hcur = hnext;
        hnext = hnn;
        kcur = knext;
    }
    __syncthreads();
    for (int64_t k = lane; k < span_len; k += 64) A.values[lo * cc + k] = span[k];
}

template <int NR, int NC, int CR, int CC>
int launch_fast(const skfa_cross_plan* P, CrossFastArgs& A, hipStream_t s) {
    const int64_t row_bytes = int64_t(P->max_row_len) * CR * CC * 8;
    const int64_t tab_bytes = int64_t(A.tab_len) * 8;
    // S lanes per row: the smallest power of two >= 2 whose R = 64 / S rows
    // fit the LDS target (or S = 64).  At least two parts share a row: rows
    // with short spans (P2 rows x P1 columns) would otherwise walk all their
    // incidences in one lane (Taylor-Hood Stokes 48^3: 0.826 against 0.859 ms)
    int S = 2;
    const int64_t target = A.lds_target;
    while (S < 64 && (64 / S) * row_bytes + tab_bytes > target) S <<= 1;
    if ((64 / S) * row_bytes + tab_bytes > kMaxWaveLds) {
        set_error("the CSR span of one row node exceeds 64 KiB of LDS");
        return kCapacity;
    }
    A.S = S;
    A.R = 64 / S;
    const int64_t waves = (P->row_nodes + A.R - 1) / A.R;
    SKFA_REQUIRE(waves <= int64_t(INT32_MAX), "too many row nodes");
    const size_t lds = size_t(A.R * row_bytes + tab_bytes);
    k_cross_fast<NR, NC, CR, CC><<<dim3(unsigned(waves)), dim3(64), lds, s>>>(A);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

// Returns kUnsupported when no fast instantiation covers the field pair.
int dispatch_fast(const skfa_cross_plan* P, CrossFastArgs& A, hipStream_t s) {
    const int nr = P->n_row, nc = P->n_col, cr = P->C_row, cc = P->C_col;
    if (nr == 4 && nc == 10 && cr == 1 && cc == 3) return launch_fast<4, 10, 1, 3>(P, A, s);
    if (nr == 10 && nc == 4 && cr == 3 && cc == 1) return launch_fast<10, 4, 3, 1>(P, A, s);
    if (nr == 4 && nc == 10 && cr == 1 && cc == 1) return launch_fast<4, 10, 1, 1>(P, A, s);
    if (nr == 10 && nc == 4 && cr == 1 && cc == 1) return launch_fast<10, 4, 1, 1>(P, A, s);
    if (nr == 4 && nc == 4 && cr == 1 && cc == 1) return launch_fast<4, 4, 1, 1>(P, A, s);
    if (nr == 10 && nc == 10 && cr == 1 && cc == 1) return launch_fast<10, 10, 1, 1>(P, A, s);
    return kUnsupported;
}

}  // namespace
}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_assemble_cross(const skfa_cross_plan* P, const double* xyz, const int32_t* vconn, const int32_t* gconn,
                        int32_t gnodes, const double* tables, int32_t nq, const double* rtab, int32_t row_kind,
                        int32_t col_kind, int32_t mode, const double* coefficient, int64_t coef_stride_e,
                        int64_t coef_stride_q, double* values, void* stream) {
    SKFA_REQUIRE(P != nullptr, "plan is null");
    SKFA_REQUIRE(values != nullptr || P->nnz() == 0, "values pointer is null");
    SKFA_REQUIRE(P->n_row <= 10 && P->n_col <= 10, "cross assembly supports 4- and 10-node tetrahedral fields");
    SKFA_REQUIRE(row_kind == SKFA_CROSS_VALUE || row_kind == SKFA_CROSS_GRADIENT, "row kind must be value or gradient");
    SKFA_REQUIRE(col_kind == SKFA_CROSS_VALUE || col_kind == SKFA_CROSS_GRADIENT,
                 "column kind must be value or gradient");
    SKFA_REQUIRE(mode >= SKFA_CROSS_SCALAR && mode <= SKFA_CROSS_TENSOR, "unknown cross mode");
    SKFA_REQUIRE(coef_stride_e >= 0 && coef_stride_q >= 0, "negative coefficient stride");
    if (mode == SKFA_CROSS_SCALAR) {
        SKFA_REQUIRE(P->C_row == P->C_col, "scalar cross coefficient requires equal component counts");
        SKFA_REQUIRE(row_kind == col_kind, "value-gradient coupling requires a tensor coefficient");
    } else if (mode == SKFA_CROSS_ROW_DIVERGENCE) {
        SKFA_REQUIRE(P->C_row == 3 && P->C_col == 1, "row divergence requires vector test and scalar trial");
        SKFA_REQUIRE(row_kind == SKFA_CROSS_GRADIENT && col_kind == SKFA_CROSS_VALUE,
                     "row divergence couples row gradients with column values");
    } else if (mode == SKFA_CROSS_COLUMN_DIVERGENCE) {
        SKFA_REQUIRE(P->C_row == 1 && P->C_col == 3, "column divergence requires scalar test and vector trial");
        SKFA_REQUIRE(row_kind == SKFA_CROSS_VALUE && col_kind == SKFA_CROSS_GRADIENT,
                     "column divergence couples row values with column gradients");
    } else {
        SKFA_REQUIRE(coefficient != nullptr, "tensor coefficient is null");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->nnz() == 0 || P->E == 0 || P->row_nodes == 0) {
        if (P->nnz()) SKFA_HIP(hipMemsetAsync(values, 0, sizeof(double) * P->nnz(), s));
        return kOk;
    }
    SKFA_REQUIRE(xyz != nullptr && vconn != nullptr && tables != nullptr, "geometry pointer is null");
    SKFA_REQUIRE(nq >= 1 && nq <= 4096, "quadrature size out of range");
    SKFA_REQUIRE(gconn == nullptr || (gnodes >= 4 && gnodes <= 64), "geometry node count out of range");
    // affine elements with element-constant scalar coefficients take the
    // reduced-table kernel where an instantiation covers the field pair
    if (rtab != nullptr && gconn == nullptr && mode != SKFA_CROSS_TENSOR && coef_stride_q == 0) {
        CrossFastArgs F{};
        F.inc_ptr = P->inc_ptr.p;
        F.nptr = P->indptr.p;
        F.records = P->records.p;
        F.rec_bytes = P->rec_bytes;
        F.off_bytes = P->off_bytes;
        F.row_nodes = P->row_nodes;
        F.xyz = reinterpret_cast<const double4*>(xyz);
        F.vconn = vconn;
        F.rtab = rtab;
        F.row_grad = row_kind == SKFA_CROSS_GRADIENT;
        F.col_grad = col_kind == SKFA_CROSS_GRADIENT;
        F.tab_len = P->n_row * P->n_col * (F.row_grad ? 3 : 1) * (F.col_grad ? 3 : 1);
        F.mode = mode;
        F.coef = coefficient;
        F.cse = coef_stride_e;
        F.values = values;
        F.lds_target = kCrossLdsTarget;
        const int rc = dispatch_fast(P, F, s);
        if (rc != kUnsupported) return rc;
    }
    // rows per wave: lane groups of the next power of two above the entry
    // count of one element row, fewer rows when their spans outgrow LDS
    const int nent = P->C_row * P->n_col * P->C_col;
    int group = 1;
    while (group < nent && group < 64) group <<= 1;
    int rows = 64 / group;
    const int64_t row_bytes = int64_t(P->max_row_len) * P->C_row * P->C_col * 8;
    while (rows > 1 && rows * row_bytes > kMaxWaveLds) rows >>= 1;
    if (row_bytes > kMaxWaveLds) {
        set_error("the CSR span of one row node exceeds 64 KiB of LDS");
        return kCapacity;
    }
    CrossArgs A{};
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
    A.xyz = reinterpret_cast<const double4*>(xyz);
    A.vconn = vconn;
    A.gconn = gconn;
    A.gn = gconn ? gnodes : 4;
    A.tables = tables;
    A.Q = nq;
    A.row_grad = row_kind == SKFA_CROSS_GRADIENT;
    A.col_grad = col_kind == SKFA_CROSS_GRADIENT;
    A.mode = mode;
    A.coef = coefficient;
    A.cse = coef_stride_e;
    A.csq = coef_stride_q;
    A.values = values;
    const int64_t waves = (P->row_nodes + rows - 1) / rows;
    SKFA_REQUIRE(waves <= int64_t(INT32_MAX), "too many row nodes");
    const size_t lds = size_t(rows * row_bytes);
    k_cross_rows<<<dim3(unsigned(waves)), dim3(64), lds, s>>>(A);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

}  // extern "C"
