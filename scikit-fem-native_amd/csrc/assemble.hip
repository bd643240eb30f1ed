// Row-gather element assembly for gfx950 (MI355X).
//
// Replaces the element loops of the reference's CPU assemblers:
//  * BilinearFormAssembler::assemble / assemble_element
//    (cpp/src/assembly/bilinear_form_assembler.cpp:39-105): value, gradient,
//    symmetric-gradient and divergence kinds over scalar and vector fields;
//  * CrossBilinearAssembler::assemble, gradient x gradient tensor path
//    (cpp/src/assembly/cross_bilinear_assembler.cpp:55-107,207-247,
//     cpp/include/native_fem/cross_contraction.hpp:42-84);
//  * tabulate_basis_geometry (cpp/src/assembly/basis_geometry.cpp:49-209),
//    here fused into the element kernel (J, det, adjugate per element or per
//    quadrature point) and also available as a standalone tabulation;
//  * integrate_tet_n1 + the np.add.at scatter
//    (cpp/src/assembly/edge_topology.cpp:156-233, hcurl_assembler.py:94-97);
//  * LinearFormAssembler (cpp/src/assembly/linear_form_assembler.cpp:41-137).
//
// Execution model.  One 64-lane wave owns a contiguous range of CSR node
// rows (64 rows for scalar fields, 7 node rows x 9 component lanes for 3D
// vector fields).  The CSR values of that range form one contiguous span;
// the wave accumulates it in LDS, each lane owning the entries of its own
// (row, column-component) so no two lanes ever touch one address.  Lanes
// step in lockstep through their row's element incidences (the gather plan
// built in pattern.hip): per step each lane recomputes the geometry of one
// element, evaluates one row of that element's matrix in reference
// coordinates and adds it at the precomputed in-row offsets.  The span is
// then written to HBM once, coalesced.  The summation order per entry is
// fixed by the plan (elements in ascending order), so results are
// bit-reproducible run to run; there are no global atomics and no colouring
// (the span adds are LDS ds_add_f64 with one writing lane per address).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "skfa_internal.h"
#include "skfa_plan.h"
#include "../../include/skfem_amd.h"

namespace skfa {

// Layout of the reference tables buffer (FP64), built by the host:
//   w[Q] | phi[Q][n] | dphi[Q][n][3] | gphi[Q][gn] | gdphi[Q][gn][3]
// For N1 the "phi" block holds vector values phi[Q][6][3] (dphi absent) and
// a trailing curl[6][3] block.
struct Tables {
    const double* w;
    const double* phi;
    const double* dphi;
    const double* gphi;
    const double* gdphi;
};

__device__ inline Tables tables_at(const double* t, int Q, int n, int gn) {
    Tables T;
    T.w = t;
    T.phi = T.w + Q;
    T.dphi = T.phi + Q * n;
    T.gphi = T.dphi + Q * n * 3;
    T.gdphi = T.gphi + Q * gn;
    return T;
}

struct Geo {
    double adj[9];  // adjugate of J (J^{-1} = adj / det), row-major
    double det;
};

__device__ inline void adjugate(const double* J, Geo& g) {
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

// Affine map from the four corner vertices: J[:,k] = v_{k+1} - v_0.
__device__ inline void affine_jacobian(const double4* __restrict__ xyz, const int32_t* __restrict__ c4, double* J) {
    const double4 v0 = xyz[c4[0]], v1 = xyz[c4[1]], v2 = xyz[c4[2]], v3 = xyz[c4[3]];
    J[0] = v1.x - v0.x; J[1] = v2.x - v0.x; J[2] = v3.x - v0.x;
    J[3] = v1.y - v0.y; J[4] = v2.y - v0.y; J[5] = v3.y - v0.y;
    J[6] = v1.z - v0.z; J[7] = v2.z - v0.z; J[8] = v3.z - v0.z;
}

// Isoparametric map: J_ij = sum_a x_{a,i} dN_a/dxi_j at quadrature point q
// (basis_geometry.cpp:112-129).
__device__ inline void iso_jacobian(const double4* __restrict__ xyz, const int32_t* __restrict__ cg, int gn,
                                    const double* __restrict__ gd, double* J) {
#pragma unroll
    for (int k = 0; k < 9; ++k) J[k] = 0.;
    for (int a = 0; a < gn; ++a) {
        const double4 x = xyz[cg[a]];
        const double g0 = gd[a * 3 + 0], g1 = gd[a * 3 + 1], g2 = gd[a * 3 + 2];
        J[0] += x.x * g0; J[1] += x.x * g1; J[2] += x.x * g2;
        J[3] += x.y * g0; J[4] += x.y * g1; J[5] += x.y * g2;
        J[6] += x.z * g0; J[7] += x.z * g1; J[8] += x.z * g2;
    }
}

// Coefficients of one element (or one quadrature point).
struct Coef {
    double m, g, s, lam;
    double A[9];
};

// ------------------------------------------------------------------ plan view
struct PlanArgs {
    const int64_t* indptr;     // node-level CSR row pointers
    const int64_t* wave_r0;    // nwaves+1
    const int64_t* wave_rec;   // nwaves+1
    const int32_t* wave_split; // nwaves
    const uint8_t* records;
    int64_t nnodes;
    int64_t nwaves;
    const int64_t* xcd_seg;  // 9 wave offsets (placement mode 1)
    const WaveMeta* meta;    // per-wave constants
    int32_t placement;
    int32_t xcd_chunk;
    int32_t streaming;       // non-temporal record loads and value stores
    int32_t diag3 = 0;       // scalar plan written as the 3x3 diagonal blocks of a 3-component field
    int32_t concurrent = 0;  // row parts add in the same instructions (row split 3 with ds_add)
};

// Lane role inside a wave: node row r0+i, components (c,d) of the CxC node
// block, part p of the row's S parts (how parts share a row: SP_* below).
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8).  Wave placement modes (plan tuning):
//   0: plain order (wave = block);
//   1: XCD x owns the contiguous wave segment [seg[x], seg[x+1]) of about
//      1/8 of the gather work, so each XCD's L2 holds only its part of the
//      mesh; the grid is 8 * max segment and surplus blocks exit at once;
//   2: chunks of `chunk` consecutive waves dealt round-robin to the XCDs.
// Returns -1 for a surplus block.
__device__ inline int placed_wave(const int64_t* seg, int nwaves, int mode, int chunk) {
    const int b = blockIdx.x;
    if (mode == 1) {
        const int x = b & 7, k = b >> 3;
        const int64_t w = seg[x] + k;
        return w < seg[x + 1] ? int(w) : -1;
    }
    if (mode == 2 && chunk > 0) {
        const int x = b & 7, k = b >> 3;
        const int w = ((k / chunk) * 8 + x) * chunk + (k % chunk);
        const int full = (nwaves / (8 * chunk)) * (8 * chunk);
        return b < full ? w : b;
    }
    return b;
}

// steps counts the wave's incidence layers; a lane's k-th step reads layer
// k * lmul + ladd (lmul = 1, ladd = 0 unless the row's lanes split incidences)
struct Lane {
    int i, c, d, p, S, R;
    int lmul, ladd;
    bool active;
    int64_t rec0, steps, lo, span, rb;
};

template <int C>
__device__ inline void lane_setup(const PlanArgs& P, int w, int lane, Lane& L) {
    const WaveMeta m = P.meta[w];
    const int64_t r0 = m.r0;
    L.R = m.R;
    L.S = m.S;
    const int lpr = C * C * L.S;
    L.i = lane / lpr;
    const int rem = lane - L.i * lpr;
    const int cd = rem / L.S;
    L.p = rem & (L.S - 1);
    L.c = cd / C;
    L.d = cd - L.c * C;
    L.active = L.i < L.R;
    L.lmul = 1;
    L.ladd = 0;
    L.rec0 = m.rec0;
    L.steps = m.steps;
    L.lo = m.lo;
    L.span = int64_t(m.span) * C * C;
    L.rb = 0;
    if (L.active) {
        const int64_t rn = r0 + L.i;
        const int64_t b = P.indptr[rn];
        const int64_t len = P.indptr[rn + 1] - b;
        L.rb = int64_t(C) * C * (b - L.lo) + int64_t(L.c) * C * len + L.d;
    }
}

// Span accumulation: one ds_add_f64 (no return value) per entry.  Outside
// the concurrent row split every LDS address has one writing lane per
// instruction, so the atomic only saves the read-back (one LDS instruction
// instead of ds_read + v_add + ds_write) and the per-entry summation order is
// the lanes' program order.  The add is unconditional: a runtime choice of
// accumulation put a uniform branch around every add and kept the compiler
// from grouping a step's adds.
__device__ inline void lds_add(double* p, double v) { (void)atomicAdd(p, v); }

typedef double f64x2 __attribute__((ext_vector_type(2)));

// The span starts at the (16-byte aligned) base of the dynamic LDS: zero it
// two doubles per lane and instruction.
__device__ inline void zero_span(double* lds, int64_t span, int lane) {
    f64x2* l2 = reinterpret_cast<f64x2*>(lds);
    const int64_t npair = span >> 1;
    for (int64_t k = lane; k < npair; k += 64) l2[k] = f64x2{0., 0.};
    if ((span & 1) && lane == 0) lds[span - 1] = 0.;
}

// Span -> HBM with 16-byte stores: one leading element when `out` is not
// 16-byte aligned, then pairs, then a trailing element.
__device__ inline void write_span(const double* lds, int64_t span, int lane, double* out, bool streaming = false) {
    if (span <= 0) return;
    const int head = (reinterpret_cast<uintptr_t>(out) & 15) ? 1 : 0;
    const int64_t npair = (span - head) >> 1;
    f64x2* o2 = reinterpret_cast<f64x2*>(out + head);
    const double* l = lds + head;
    if (streaming) {
        if (head && lane == 0) __builtin_nontemporal_store(lds[0], out);
        for (int64_t k = lane; k < npair; k += 64) {
            const f64x2 v = {l[2 * k], l[2 * k + 1]};
            __builtin_nontemporal_store(v, o2 + k);
        }
        if (((span - head) & 1) && lane == 0) __builtin_nontemporal_store(lds[span - 1], out + span - 1);
    } else {
        if (head && lane == 0) out[0] = lds[0];
        for (int64_t k = lane; k < npair; k += 64) o2[k] = f64x2{l[2 * k], l[2 * k + 1]};
        if (((span - head) & 1) && lane == 0) out[span - 1] = lds[span - 1];
    }
}

// Span -> HBM through a map: entry k goes to out[map[k]] unless map[k] < 0
// (map points at the span's first entry).  Kept entries keep their order,
// so a wave's stores stay nearly contiguous.
__device__ inline void write_span_mapped(const double* lds, int64_t span, int lane, const int32_t* __restrict__ map,
                                         double* out, bool streaming) {
    for (int64_t k = lane; k < span; k += 64) {
        const int32_t j = map[k];
        if (j < 0) continue;
        if (streaming)
            __builtin_nontemporal_store(lds[k], out + j);
        else
            out[j] = lds[k];
    }
}

// Write a scalar span as the 3x3 node blocks K_(a,c),(b,d) = [c == d] s_ab
// of a 3-component field (value / gradient kinds of vector fields): node row
// r with len columns becomes DOF rows 3r+c, whose entries (column node j,
// component d) sit at 9 nptr[r] + 3 c len + 3 j + d.
__device__ inline void write_span_diag3(const double* lds, const WaveMeta& m, const int64_t* __restrict__ nptr,
                                        int lane, double* values, bool streaming) {
    for (int i = 0; i < m.R; ++i) {
        const int64_t b = nptr[m.r0 + i];
        const int len = int(nptr[m.r0 + i + 1] - b);
        const double* sr = lds + (b - m.lo);
        double* out = values + 9 * b;
        const int n = 9 * len, len3 = 3 * len;
        for (int k = lane; k < n; k += 64) {
            const int c = k / len3;
            const int rem = k - c * len3;
            const int j = rem / 3;
            const int d = rem - 3 * j;
            const double v = c == d ? sr[j] : 0.;
            if (streaming)
                __builtin_nontemporal_store(v, out + k);
            else
                out[k] = v;
        }
    }
}

template <int RB>
struct Record {
    int32_t packed;
    uint8_t off[RB - 4];
};

template <int RB>
__device__ inline void load_record(const uint8_t* __restrict__ base, Record<RB>& r, bool streaming = false) {
    if constexpr (RB == 16) {
        // records are read once per assembly: a non-temporal load keeps them
        // from evicting the reused element payloads from L2
        uint4 v;
        if (streaming) {
            const uint32_t* q = reinterpret_cast<const uint32_t*>(base);
            v.x = __builtin_nontemporal_load(q);
            v.y = __builtin_nontemporal_load(q + 1);
            v.z = __builtin_nontemporal_load(q + 2);
            v.w = __builtin_nontemporal_load(q + 3);
        } else {
            v = *reinterpret_cast<const uint4*>(base);
        }
        r.packed = int32_t(v.x);
        uint32_t* o = reinterpret_cast<uint32_t*>(r.off);
        o[0] = v.y; o[1] = v.z; o[2] = v.w;
    } else if constexpr (RB == 8) {
        const uint2 v = *reinterpret_cast<const uint2*>(base);
        r.packed = int32_t(v.x);
        *reinterpret_cast<uint32_t*>(r.off) = v.y;
    } else if constexpr (RB == 16) {
        const uint4 v = *reinterpret_cast<const uint4*>(base);
        r.packed = int32_t(v.x);
        uint32_t* o = reinterpret_cast<uint32_t*>(r.off);
        o[0] = v.y; o[1] = v.z; o[2] = v.w;
    } else {
        const uint4* q = reinterpret_cast<const uint4*>(base);
        uint32_t* o = reinterpret_cast<uint32_t*>(r.off);
        const uint4 v0 = q[0];
        r.packed = int32_t(v0.x);
        o[0] = v0.y; o[1] = v0.z; o[2] = v0.w;
#pragma unroll
        for (int k = 1; k < RB / 16; ++k) {
            const uint4 v = q[k];
            o[4 * k - 1] = v.x; o[4 * k] = v.y; o[4 * k + 1] = v.z; o[4 * k + 2] = v.w;
        }
    }
}

template <int RB, int OFFW>
__device__ inline int rec_offset(const Record<RB>& r, int b) {
    if constexpr (OFFW == 1) return r.off[b];
    else return reinterpret_cast<const uint16_t*>(r.off)[b];
}

// In-row offset of local node b (runtime b) without dynamic register indexing.
// The words are copied out one by one at compile-time positions: a runtime
// index into the record's byte array would put the record in scratch memory.
template <int RB, int OFFW>
__device__ inline int rec_offset_dyn(const Record<RB>& r, int b) {
    constexpr int NW = (RB - 4) / 4;
    constexpr int per = 4 / OFFW;   // offsets per word
    const int wi = b / per;
    uint32_t word = 0;
#pragma unroll
    for (int t = 0; t < NW; ++t) {
        uint32_t wt;
        __builtin_memcpy(&wt, r.off + 4 * t, 4);
        if (wi == t) word = wt;
    }
    const int sh = (b - wi * per) * 8 * OFFW;
    return int((word >> sh) & (OFFW == 1 ? 0xffu : 0xffffu));
}

// ------------------------------------------------------------------ element payload
// Fast path (affine elements, coefficients constant per element): one pass
// over elements stores what every element row needs, so a gather step is
// record -> payload (two dependent loads) instead of record -> connectivity
// -> coordinates/coefficients.
//   C == 1 (stride 10): D = adj X adj^T / |det| with
//            X = (g + s/2) I + (s/2 + lam) e0 e0^T + A, then |det| m
//   C == 3 (stride 14): B = adj / sqrt|det|, g + s/2, s/2, lam, |det| m
//            (lane (c,d) builds D_cd = dcd (g+s/2) B B^T + s/2 B_d B_c^T + lam B_c B_d^T);
//            the FP32-products vector-P2 gather reads the same 13 values as
//            16 floats (64 B), which it rounds to FP32 anyway
//   N1     (stride 14): kappa adj adj^T/|det| (6), wsum mu J^T J/|det| (6),
//            edge signs (bits)

struct PayloadArgs {
    const double4* xyz;
    const int32_t* vconn;
    int64_t E;
    const double *cm, *cg, *cs, *cl, *cA;
    const double* wtab;  // N1: quadrature weights (their sum scales the curl term)
    int32_t Q;
    double* out;
    // element tables of the table gather (built by block 0 when ntab is set)
    const double* rtab;
    int32_t tab_n, tab_qg, tab_qm;
    double* ntab;
    float* ntab32;
};

// Element-table entry t = a * n + b of n-node elements from the reduced
// quadrature tables (wg | dg | wm | pm): the six D slots and the mass, FP64
// and FP32.  kNK values per entry (defined with the table gather below).
__device__ inline void build_table_entry(const double* __restrict__ rtab, int n, int32_t Qg, int32_t Qm, int t,
                                         double* __restrict__ out, float* __restrict__ out32);

// Storage of the symmetric scalar payload (C == 1, SYM): 8 doubles (D upper
// triangle, |det| m, pad), 6 doubles (D only: forms without a mass term) or
// 8 floats (D, |det| m, pad; the FP32-products gather rounds D to FP32 anyway)
constexpr int SYMOUT_F64 = 0, SYMOUT_F64_NOMASS = 1, SYMOUT_F32 = 2;

// Payloads are written in blocks of kPayloadBlock consecutive elements: each
// thread puts its element's values into an LDS block buffer, and the block's
// contiguous output range then goes out as 16-byte stores.  (A lane-per-
// element store of a 48-112 B payload touches a different cache line per
// lane and value.)  Every layout is a whole number of 16-byte units and the
// blocks start at multiples of 256 elements, so the ranges are 16-byte aligned.
constexpr int kPayloadBlock = 256;

template <typename V, int PO>
__device__ inline void flush_payload_block(const V* buf, V* out, int64_t e0, int n) {
    static_assert((PO * sizeof(V)) % 16 == 0, "payload rows are whole 16-byte units");
    const int nvec = n * PO * int(sizeof(V)) / 16;
    const uint4* src = reinterpret_cast<const uint4*>(buf);
    uint4* dst = reinterpret_cast<uint4*>(out + e0 * PO);
    for (int k = threadIdx.x; k < nvec; k += blockDim.x) dst[k] = src[k];
}

template <int C, bool SYM, int SYMOUT>
struct PayloadOut {
    using V = typename std::conditional<SYMOUT == SYMOUT_F32 && (C == 3 || SYM), float, double>::type;
    static constexpr int PO = C == 1 ? (SYM ? (SYMOUT == SYMOUT_F64_NOMASS ? 6 : 8) : 10)
                                     : (SYMOUT == SYMOUT_F32 ? 16 : 14);
};

template <int C, bool SYM, int SYMOUT = SYMOUT_F64>
__global__ void __launch_bounds__(kPayloadBlock) k_payload_h1(PayloadArgs A) {
    using V = typename PayloadOut<C, SYM, SYMOUT>::V;
    constexpr int PO = PayloadOut<C, SYM, SYMOUT>::PO;
    // one LDS area: first the block's tensor rows (C == 1), read with
    // coalesced 16-byte loads instead of nine 8-byte loads per lane at a
    // 72-byte lane stride, then the block's payload rows
    constexpr int kTab = C == 1 ? kPayloadBlock * 9 * 8 : 0;
    constexpr int kBuf = kPayloadBlock * PO * int(sizeof(V));
    __shared__ __align__(16) unsigned char smem[kTab > kBuf ? kTab : kBuf];
    V* buf = reinterpret_cast<V*>(smem);
    const bool stage = C == 1 && A.cA != nullptr && (reinterpret_cast<uintptr_t>(A.cA) & 15) == 0;
    // the table gather's element tables ride along in block 0 (no kernel of
    // their own); the gather runs after this kernel in stream order
    if (A.ntab != nullptr && blockIdx.x == 0)
        for (int t = threadIdx.x; t < A.tab_n * A.tab_n; t += kPayloadBlock)
            build_table_entry(A.rtab, A.tab_n, A.tab_qg, A.tab_qm, t, A.ntab, A.ntab32);
    for (int64_t e0 = int64_t(blockIdx.x) * kPayloadBlock; e0 < A.E; e0 += int64_t(gridDim.x) * kPayloadBlock) {
        const int n = int(A.E - e0 < kPayloadBlock ? A.E - e0 : kPayloadBlock);
        double Xs[9];
        if constexpr (C == 1) {
            if (stage) {
                // e0 is a multiple of 256, so the slab starts 16-byte aligned
                double* tb = reinterpret_cast<double*>(smem);
                const double2* src = reinterpret_cast<const double2*>(A.cA + e0 * 9);
                const int nd = n * 9;
                for (int k = threadIdx.x; k < nd / 2; k += kPayloadBlock)
                    reinterpret_cast<double2*>(tb)[k] = src[k];
                if ((nd & 1) && threadIdx.x == 0) tb[nd - 1] = A.cA[e0 * 9 + nd - 1];
                __syncthreads();
                if (int(threadIdx.x) < n) {
#pragma unroll
                    for (int t = 0; t < 9; ++t) Xs[t] = tb[threadIdx.x * 9 + t];
                }
                __syncthreads();   // the payload rows reuse the area
            }
        }
        if (int(threadIdx.x) < n) {
            const int64_t e = e0 + threadIdx.x;
            V* o = buf + threadIdx.x * PO;
            double J[9];
            Geo G;
            {
                const int4 c = reinterpret_cast<const int4*>(A.vconn)[e];
                const int32_t c4[4] = {c.x, c.y, c.z, c.w};
                affine_jacobian(A.xyz, c4, J);
            }
            adjugate(J, G);
            const double ad = fabs(G.det);
            const double m = A.cm ? A.cm[e] : 0.;
            const double g = A.cg ? A.cg[e] : 0.;
            const double s = A.cs ? A.cs[e] : 0.;
            const double lam = A.cl ? A.cl[e] : 0.;
            const double* a = G.adj;
            if constexpr (C == 1) {
                double X[9];
#pragma unroll
                for (int t = 0; t < 9; ++t) X[t] = stage ? Xs[t] : (A.cA ? A.cA[e * 9 + t] : 0.);
                X[0] += g + 0.5 * s + 0.5 * s + lam;
                X[4] += g + 0.5 * s;
                X[8] += g + 0.5 * s;
                double T[9];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        T[i * 3 + j] = a[i * 3 + 0] * X[0 * 3 + j] + a[i * 3 + 1] * X[1 * 3 + j] + a[i * 3 + 2] * X[2 * 3 + j];
                const double inv = 1. / ad;
                double Dm[9];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        Dm[i * 3 + j] = inv * (T[i * 3 + 0] * a[j * 3 + 0] + T[i * 3 + 1] * a[j * 3 + 1] + T[i * 3 + 2] * a[j * 3 + 2]);
                if constexpr (SYM) {
                    // X symmetric => D symmetric; store the upper triangle
                    o[0] = V(Dm[0]); o[1] = V(Dm[1]); o[2] = V(Dm[2]);
                    o[3] = V(Dm[4]); o[4] = V(Dm[5]); o[5] = V(Dm[8]);
                    if constexpr (PO == 8) {
                        o[6] = V(ad * m);
                        o[7] = V(ad);   // |det J|: the load vector fused into the element-table gather
                    }
                } else {
#pragma unroll
                    for (int t = 0; t < 9; ++t) o[t] = Dm[t];
                    o[9] = ad * m;
                }
            } else {
                const double sc = 1. / sqrt(ad);
#pragma unroll
                for (int t = 0; t < 9; ++t) o[t] = V(a[t] * sc);
                o[9] = V(g + 0.5 * s);
                o[10] = V(0.5 * s);
                o[11] = V(lam);
                o[12] = V(ad * m);
#pragma unroll
                for (int t = 13; t < PO; ++t) o[t] = V(0);
            }
        }
        __syncthreads();
        flush_payload_block<V, PO>(buf, reinterpret_cast<V*>(A.out), e0, n);
        __syncthreads();
    }
}

__constant__ int kN1Edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

__global__ void __launch_bounds__(kPayloadBlock) k_payload_n1(PayloadArgs A) {
    constexpr int PO = 14;
    __shared__ double buf[kPayloadBlock * PO];
    double wsum = 0.;
    for (int q = 0; q < A.Q; ++q) wsum += A.wtab[q];
    for (int64_t e0 = int64_t(blockIdx.x) * kPayloadBlock; e0 < A.E; e0 += int64_t(gridDim.x) * kPayloadBlock) {
        const int n = int(A.E - e0 < kPayloadBlock ? A.E - e0 : kPayloadBlock);
        if (int(threadIdx.x) < n) {
            const int64_t e = e0 + threadIdx.x;
            const int32_t* c4 = A.vconn + e * 4;
            double J[9];
            Geo G;
            affine_jacobian(A.xyz, c4, J);
            adjugate(J, G);
            const double inv = 1. / fabs(G.det);
            // element-constant coefficients are folded in; per-point ones
            // (cs/cl flags set) are applied in the gather kernel
            const double k = A.cs ? 1. : (A.cm ? A.cm[e] : 0.);
            const double mu = A.cl ? 1. : (A.cg ? A.cg[e] * wsum : 0.);
            const double* a = G.adj;
            double* o = buf + threadIdx.x * PO;
            const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                const int i = ij[t][0], j = ij[t][1];
                o[t] = k * inv * (a[i * 3 + 0] * a[j * 3 + 0] + a[i * 3 + 1] * a[j * 3 + 1] + a[i * 3 + 2] * a[j * 3 + 2]);
                o[6 + t] = mu * inv * (J[0 * 3 + i] * J[0 * 3 + j] + J[1 * 3 + i] * J[1 * 3 + j] + J[2 * 3 + i] * J[2 * 3 + j]);
            }
            int32_t v[4] = {c4[0], c4[1], c4[2], c4[3]};
            int64_t bits = 0;
#pragma unroll
            for (int l = 0; l < 6; ++l) bits |= int64_t(v[kN1Edge[l][0]] < v[kN1Edge[l][1]] ? 0 : 1) << l;
            o[12] = __longlong_as_double(bits);
            o[13] = 0.;
        }
        __syncthreads();
        flush_payload_block<double, PO>(buf, A.out, e0, n);
        __syncthreads();
    }
}

// ------------------------------------------------------------------ gather kernels
// Reference tables in LDS behind the span: gradient block wg[Qg] | dg[Qg][n][3],
// mass block wm[Qm] | pm[Qm][n].
struct GatherArgs {
    PlanArgs plan;
    const double* payload;
    const double* tab;      // FP64: wg | dg | wm | pm
    const float* tab32;     // FP32 copy
    int32_t Qg, Qm;
    int32_t has_mass;
    double* values;
    // only waves with rmin <= rows <= rmax are assembled by this launch
    int32_t rmin = 0, rmax = 1 << 30;
    // 10-node (P2) barycentric coefficients of the gradient rule: for the 16
    // (node b, barycentric index m) pairs with grad(phi_b) = sum_m c_bm(q) grad(L_m),
    // c_bm at the Qg points (pair-major); null when not available
    const double* ctab = nullptr;
    const float* ctab32 = nullptr;
    // 10-node pair table (10 x 10 x 8: q_(ab,0), q_(ab,1), M_ab, 0) of the
    // vector-P2 gather's barycentric form; null when not available
    const double* qtab = nullptr;
    const float* qtab32 = nullptr;
    // load vector fused into the element-table gather of 4-node scalar
    // fields (skfa_assemble_bilinear_load): b_i = f sum_(e,a) |det J_e| w_a,
    // w_a = sum_q W_q phi_a(X_q) from the basis tables (w | phi | ...)
    const double* load_tables = nullptr;
    int32_t load_q = 0;
    const double* load_f = nullptr;
    double* load_out = nullptr;
    // optional output maps of the fused path: entry k of the plan's CSR goes
    // to values[value_map[k]] and node row i's load to load_out[row_map[i]]
    // (-1: dropped), e.g. the Dirichlet-condensed system
    const int32_t* value_map = nullptr;
    const int32_t* row_map = nullptr;
};

// K_ab += w_q (dphi_aq^T D) . dphi_bq for all b: the lane's row a reads the
// LDS copy of the table, the b loop reads the same entries for all lanes.
template <int NL, typename T>
__device__ inline void grad_row(int q, int a, const T* D, const T* __restrict__ dg, const T* dg_l, T* Kr) {
    // dg_l holds w_q * dphi (weights folded in), dg the unweighted table
    const T* ga = dg_l + (q * NL + a) * 3;
    const T g0 = ga[0], g1 = ga[1], g2 = ga[2];
    const T u0 = fma(g0, D[0], fma(g1, D[3], g2 * D[6]));
    const T u1 = fma(g0, D[1], fma(g1, D[4], g2 * D[7]));
    const T u2 = fma(g0, D[2], fma(g1, D[5], g2 * D[8]));
#pragma unroll
    for (int b = 0; b < NL; ++b) {
        const T* gb = dg + (q * NL + b) * 3;
        Kr[b] = fma(u0, gb[0], fma(u1, gb[1], fma(u2, gb[2], Kr[b])));
    }
}

// Payload layouts of the gather kernel (values per element).
//   PL_FULL    (C==1): D (9, row-major), |det| m                 10 doubles
//   PL_SYM     (C==1, symmetric D): D00 D01 D02 D11 D12 D22, |det| m, pad
//   PL_SYM6    (C==1, symmetric D, no mass term): D00 .. D22      6 doubles
//   PL_SYM32   (C==1, symmetric D, FP32 products): as PL_SYM      8 floats
//   PL_VEC     (C==3): B (9), g + s/2, s/2, lam, |det| m, pad    14 doubles
constexpr int PL_FULL = 0, PL_SYM = 1, PL_SYM6 = 2, PL_VEC = 3, PL_SYM32 = 4;
template <int PL>
struct PayloadLayout {
    static constexpr int stride = PL == PL_FULL ? 10 : PL == PL_SYM ? 8 : PL == PL_SYM6 ? 6 : PL == PL_SYM32 ? 8 : 14;
    using type = typename std::conditional<PL == PL_SYM32, float, double>::type;
};

template <int K, int RB, class Args>
__device__ inline void load_records(const Args& A, const Lane& L, int64_t st, Record<RB>* rec) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        rec[k].packed = -1;
        const int64_t layer = (st + k) * L.lmul + L.ladd;
        if (L.active && layer < L.steps)
            load_record<RB>(A.plan.records + (L.rec0 + layer * L.R + L.i) * RB, rec[k], A.plan.streaming);
    }
}

template <int K, int RB, int P, typename PT, class Args>
__device__ inline void load_payloads(const Args& A, const Record<RB>* rec, PT (*pl)[P]) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (rec[k].packed >= 0) {
            if constexpr (sizeof(PT) == 4) {
                static_assert(P % 4 == 0, "float payloads are read as float4");
                const float4* src = reinterpret_cast<const float4*>(A.payload) + int64_t(rec[k].packed >> 4) * (P / 4);
#pragma unroll
                for (int t = 0; t < P / 4; ++t) {
                    const float4 v = src[t];
                    pl[k][4 * t] = v.x;
                    pl[k][4 * t + 1] = v.y;
                    pl[k][4 * t + 2] = v.z;
                    pl[k][4 * t + 3] = v.w;
                }
            } else {
                const double2* src = reinterpret_cast<const double2*>(A.payload + int64_t(rec[k].packed >> 4) * P);
#pragma unroll
                for (int t = 0; t < P / 2; ++t) {
                    const double2 v = src[t];
                    pl[k][2 * t] = v.x;
                    pl[k][2 * t + 1] = v.y;
                }
            }
        }
    }
}

// How the S lanes of one node row share its work (all bit-identical: every
// CSR entry receives its element contributions in incidence order):
//   SP_INCIDENCE  lane p takes incidence layers p, p+S, ...; the S lanes
//                 add their element rows in turn (p = 0, 1, ...)
//   SP_ENTRY      all S lanes read the same element; lane p computes entries
//                 b = p, p+S, ... of its row
//   SP_COLUMN     all S lanes read the same element and compute the whole
//                 row; lane p adds the in-row offsets o with o % S == p
//   SP_CONCURRENT as SP_INCIDENCE, but the S lanes add their element rows
//                 in the same instructions: the LDS atomic add serialises
//                 the lanes that hit one address (requires ds_add; the
//                 order among parts is the hardware's, so the bits may
//                 differ from the other splits by rounding)
constexpr int SP_INCIDENCE = 0, SP_ENTRY = 1, SP_COLUMN = 2, SP_CONCURRENT = 3;

// P2 element row from the barycentric structure of the reference gradients
// (the rule enters only through c_bm(q)):
//   grad(phi_b)(q) = sum_{m in supp b} c_bm(q) g_m,  g_0 = (-1,-1,-1), g_m = e_{m-1}
//   K_ab = sum_q u_q . grad(phi_b)(q) = sum_{m in supp b} sum_q c_bm(q) t_q^m,
//   t_q^m = u_q . g_m  (t^0 = -(u_0 + u_1 + u_2), t^m = u_{m-1})
// Vertex b = k has one pair (k, k); edge b = 4 + e, e = (i, j), has the
// pairs (i) and (j) at rows 4 + 2e and 5 + 2e of ctab.  4 + 6 x 2 pairs:
// 64 FMAs for the row instead of 120 (the table entries are mostly zero).
template <int QG, typename T>
__device__ inline void struct_row(const T (*u)[3], const T* __restrict__ ct, T* Kr) {
    T t[4][QG];
#pragma unroll
    for (int q = 0; q < QG; ++q) {
        t[0][q] = -((u[q][0] + u[q][1]) + u[q][2]);
        t[1][q] = u[q][0];
        t[2][q] = u[q][1];
        t[3][q] = u[q][2];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < QG; ++q) acc = fma(ct[k * QG + q], t[k][q], acc);
        Kr[k] = acc;
    }
    constexpr int E0[6] = {0, 1, 2, 0, 1, 2}, E1[6] = {1, 2, 0, 3, 3, 3};
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < QG; ++q) {
            acc = fma(ct[(4 + 2 * e) * QG + q], t[E0[e]][q], acc);
            acc = fma(ct[(5 + 2 * e) * QG + q], t[E1[e]][q], acc);
        }
        Kr[4 + e] = acc;
    }
}

// Always inlined: a call would pass the payload registers by pointer and put
// them in scratch memory (the 3-component instances were not inlined).
template <int NL, int C, int PL, bool F32, int RB, int OFFW, int QG, int SP, int KF = 0>
__device__ __forceinline__ void compute_step(const GatherArgs& A, const Lane& L, const Record<RB>& rec,
                                    const typename PayloadLayout<PL>::type* pl,
                                    double* span_lds, int Qg, int Qm, bool mass,
                                    const typename std::conditional<F32, float, double>::type* dg,
                                    const typename std::conditional<F32, float, double>::type* pm,
                                    const typename std::conditional<F32, float, double>::type* dg_l,
                                    const typename std::conditional<F32, float, double>::type* pm_l,
                                    const typename std::conditional<F32, float, double>::type* dgu_l,
                                    const typename std::conditional<F32, float, double>::type* pmu_l) {
    using T = typename std::conditional<F32, float, double>::type;
    const int a = rec.packed & 15;
    T D[9];
    T mf;
    if constexpr (PL == PL_FULL) {
#pragma unroll
        for (int t = 0; t < 9; ++t) D[t] = T(pl[t]);
        mf = T(pl[9]);
    } else if constexpr (PL == PL_SYM || PL == PL_SYM6 || PL == PL_SYM32) {
        D[0] = T(pl[0]); D[1] = T(pl[1]); D[2] = T(pl[2]);
        D[3] = T(pl[1]); D[4] = T(pl[3]); D[5] = T(pl[4]);
        D[6] = T(pl[2]); D[7] = T(pl[4]); D[8] = T(pl[5]);
        mf = PL == PL_SYM6 ? T(0) : T(pl[PL == PL_SYM6 ? 0 : 6]);
    } else {
        const double* B = pl;
        const double iso = (L.c == L.d) ? pl[9] : 0.;
        const double hs = pl[10], lam = pl[11];
        const int c = L.c, d = L.d;
        // columns c and d of B through 0/1 weights (B is finite, so the
        // sums are exact): a runtime index into the payload registers, or a
        // select between two of them, would move the payload to scratch memory
        const double wc[3] = {c == 0 ? 1. : 0., c == 1 ? 1. : 0., c == 2 ? 1. : 0.};
        const double wd[3] = {d == 0 ? 1. : 0., d == 1 ? 1. : 0., d == 2 ? 1. : 0.};
        double Bc[3], Bd[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Bc[i] = fma(B[i * 3], wc[0], fma(B[i * 3 + 1], wc[1], B[i * 3 + 2] * wc[2]));
            Bd[i] = fma(B[i * 3], wd[0], fma(B[i * 3 + 1], wd[1], B[i * 3 + 2] * wd[2]));
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double bb = B[i * 3 + 0] * B[j * 3 + 0] + B[i * 3 + 1] * B[j * 3 + 1] + B[i * 3 + 2] * B[j * 3 + 2];
                D[i * 3 + j] = T(iso * bb + hs * Bd[i] * Bc[j] + lam * Bc[i] * Bd[j]);
            }
        mf = (L.c == L.d) ? T(pl[12]) : T(0);
    }
    if constexpr (SP == SP_ENTRY) {
        // lane p computes entries b = p, p + S, ... of row a with the same
        // operation order as the full row below (bit-identical)
        auto u_of = [&](int q, T* u) {
            const T* ga = dg_l + (q * NL + a) * 3;
            const T g0 = ga[0], g1 = ga[1], g2 = ga[2];
            u[0] = fma(g0, D[0], fma(g1, D[3], g2 * D[6]));
            u[1] = fma(g0, D[1], fma(g1, D[4], g2 * D[7]));
            u[2] = fma(g0, D[2], fma(g1, D[5], g2 * D[8]));
        };
        auto finish = [&](int b, T k) {
            if (mass)
                for (int q = 0; q < Qm; ++q) k = fma(mf * pm_l[q * NL + a], pmu_l[q * NL + b], k);
            lds_add(span_lds + L.rb + int64_t(C) * rec_offset_dyn<RB, OFFW>(rec, b), double(k));
        };
        if constexpr (QG > 0) {
            T u[QG][3];
#pragma unroll
            for (int q = 0; q < QG; ++q) u_of(q, u[q]);
            for (int b = L.p; b < NL; b += L.S) {
                T k = T(0);
#pragma unroll
                for (int q = 0; q < QG; ++q) {
                    const T* gb = dgu_l + (q * NL + b) * 3;
                    k = fma(u[q][0], gb[0], fma(u[q][1], gb[1], fma(u[q][2], gb[2], k)));
                }
                finish(b, k);
            }
        } else {
            for (int b = L.p; b < NL; b += L.S) {
                T k = T(0);
                for (int q = 0; q < Qg; ++q) {
                    T u[3];
                    u_of(q, u);
                    const T* gb = dgu_l + (q * NL + b) * 3;
                    k = fma(u[0], gb[0], fma(u[1], gb[1], fma(u[2], gb[2], k)));
                }
                finish(b, k);
            }
        }
        return;
    }
    T Kr[NL];
#pragma unroll
    for (int b = 0; b < NL; ++b) Kr[b] = T(0);
    if constexpr (KF == 1 && NL == 10 && QG > 0) {
        // u_q = D^T (w_q grad phi_a(q)), then the structural row
        T u[QG][3];
#pragma unroll
        for (int q = 0; q < QG; ++q) {
            const T* ga = dg_l + (q * NL + a) * 3;
            const T g0 = ga[0], g1 = ga[1], g2 = ga[2];
            u[q][0] = fma(g0, D[0], fma(g1, D[3], g2 * D[6]));
            u[q][1] = fma(g0, D[1], fma(g1, D[4], g2 * D[7]));
            u[q][2] = fma(g0, D[2], fma(g1, D[5], g2 * D[8]));
        }
        const T* ct = F32 ? reinterpret_cast<const T*>(A.ctab32) : reinterpret_cast<const T*>(A.ctab);
        struct_row<QG, T>(u, ct, Kr);
    } else if constexpr (QG > 0) {
#pragma unroll
        for (int q = 0; q < QG; ++q) grad_row<NL, T>(q, a, D, dg, dg_l, Kr);
    } else {
        for (int q = 0; q < Qg; ++q) grad_row<NL, T>(q, a, D, dg, dg_l, Kr);
    }
    if (mass) {
        for (int q = 0; q < Qm; ++q) {
            const T v = mf * pm_l[q * NL + a];  // pm_l holds w_q * phi
#pragma unroll
            for (int b = 0; b < NL; ++b) Kr[b] = fma(v, pm[q * NL + b], Kr[b]);
        }
    }
    if constexpr (SP == SP_COLUMN) {
#pragma unroll
        for (int b = 0; b < NL; ++b) {
            const int o = rec_offset<RB, OFFW>(rec, b);
            if ((o & (L.S - 1)) == L.p) lds_add(span_lds + L.rb + int64_t(C) * o, double(Kr[b]));
        }
    } else {
        // LDS addresses fit 32 bits; one writer per address in every add
        double* row = span_lds + int(L.rb);
        auto add_row = [&]() {
#pragma unroll
            for (int b = 0; b < NL; ++b)
                lds_add(row + C * rec_offset<RB, OFFW>(rec, b), double(Kr[b]));
        };
        if (L.S == 1 || SP == SP_CONCURRENT) {
            add_row();
        } else {
            for (int turn = 0; turn < L.S; ++turn)
                if (L.p == turn) add_row();
        }
    }
}

// Software-pipelined gather: while batch b is computed, the payloads of
// batch b+1 and the records of batch b+2 are in flight, so the two
// dependent loads of a step (record -> payload) overlap other work.  (The
// vector-P2 gather below swaps two buffer sets instead of copying; here
// that measured slower: C2 0.264 -> 0.273 ms, C4 0.699 -> 0.707 ms.)
template <int NL, int C, int PL, bool F32, int RB, int OFFW, int QG, int SP, int KF = 0>
__global__ void __launch_bounds__(64) k_gather_h1(GatherArgs A) {
    extern __shared__ double span_lds[];
    constexpr int P = PayloadLayout<PL>::stride;
    constexpr int K = 2;
    constexpr bool ES = SP == SP_ENTRY;
    using T = typename std::conditional<F32, float, double>::type;
    const int lane = threadIdx.x;
    Lane L;
    const int wv = placed_wave(A.plan.xcd_seg, int(A.plan.nwaves), A.plan.placement, A.plan.xcd_chunk);
    if (wv < 0) return;
    lane_setup<C>(A.plan, wv, lane, L);
    if (L.R < A.rmin || L.R > A.rmax) return;
    int64_t nsteps = L.steps;
    if constexpr (SP == SP_INCIDENCE || SP == SP_CONCURRENT) {
        L.lmul = L.S;
        L.ladd = L.p;
        nsteps = (L.steps + L.S - 1) / L.S;
    }
    using PT = typename PayloadLayout<PL>::type;
    Record<RB> r0[K], r1[K], r2[K];
    PT p0[K][P], p1[K][P];
    load_records<K, RB>(A, L, 0, r0);
    zero_span(span_lds, L.span, lane);
    // stage the lane-varying tables (rows a of dg and pm) behind the span
    const int Qg = QG > 0 ? QG : A.Qg;  // compile-time for the 1- and 4-point reduced rules
    const int Qm = A.Qm;
    const int tab_len = Qg + Qg * NL * 3 + Qm + Qm * NL;
    T* tab_lds = reinterpret_cast<T*>(span_lds + L.span);
    const T* tab_src = F32 ? reinterpret_cast<const T*>(A.tab32) : reinterpret_cast<const T*>(A.tab);
    for (int k = lane; k < tab_len; k += 64) {
        T v = tab_src[k];
        if (k >= Qg && k < Qg + Qg * NL * 3) v *= tab_src[(k - Qg) / (NL * 3)];  // w_q * dphi
        if (k >= Qg + Qg * NL * 3 + Qm) v *= tab_src[Qg + Qg * NL * 3 + (k - Qg - Qg * NL * 3 - Qm) / NL];  // w_q * phi
        tab_lds[k] = v;
    }
    // entry split: the unweighted tables are read at lane-varying b too
    T* tabu_lds = tab_lds + tab_len;
    if constexpr (ES) {
        const int lg = Qg * NL * 3;
        for (int k = lane; k < lg + Qm * NL; k += 64)
            tabu_lds[k] = k < lg ? tab_src[Qg + k] : tab_src[Qg + lg + Qm + (k - lg)];
    }
    load_payloads<K, RB, P>(A, r0, p0);
    load_records<K, RB>(A, L, K, r1);
    __syncthreads();
    const T* dg = tab_src + Qg;            // uniform reads (scalar loads)
    const T* pm = dg + Qg * NL * 3 + Qm;
    const T* dg_l = tab_lds + Qg;          // lane-varying reads
    const T* pm_l = tab_lds + Qg + Qg * NL * 3 + Qm;
    const T* dgu_l = tabu_lds;
    const T* pmu_l = tabu_lds + Qg * NL * 3;
    const bool mass = A.has_mass != 0;
    for (int64_t st = 0; st < nsteps; st += K) {
        load_payloads<K, RB, P>(A, r1, p1);
        load_records<K, RB>(A, L, st + 2 * K, r2);
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (r0[k].packed >= 0)
                compute_step<NL, C, PL, F32, RB, OFFW, QG, SP, KF>(A, L, r0[k], p0[k], span_lds, Qg, Qm, mass, dg,
                                                                   pm, dg_l, pm_l, dgu_l, pmu_l);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            r0[k] = r1[k];
            r1[k] = r2[k];
#pragma unroll
            for (int t = 0; t < P; ++t) p0[k][t] = p1[k][t];
        }
    }
    __syncthreads();
    if (C == 1 && A.plan.diag3)
        write_span_diag3(span_lds, A.plan.meta[wv], A.plan.indptr, lane, A.values, A.plan.streaming);
    else
        write_span(span_lds, L.span, lane, A.values + L.lo * C * C, A.plan.streaming);
}

// ------------------------------------------------------------------ vector P2 gather
// 10-node elements, 3 components, element-constant coefficients (payload
// PL_VEC: B = adj / sqrt|det|, iso = g + s/2, hs = s/2, lam, |det| m).
// Every (c, d) block entry of the node pair (a, b) derives from one 3x3
// matrix of the pair:
//   G_ab = sum_q w_q (B^T dphi_a(q)) (B^T dphi_b(q))^T
//   K_(a,c),(b,d) = lam G_cd + hs G_dc + [c == d] (iso tr G_ab + |det| m M_ab)
// (the form of the lane-per-(c,d) gather's D_cd, contracted once per pair).
// Two ways to form G_ab (KF, the row form):
//   KF = 0  the quadrature sum above, 4 points: B^T dphi_b(q) and
//           B^T (w_q dphi_a(q)) per point, then their outer products;
//   KF = 1  B is constant on the element, so G_ab = B^T N_ab B with the
//           reference sum N_ab = sum_q w_q dphi_a dphi_b^T, and through the
//           barycentric structure of the P2 gradients N_ab has rank two:
//             N_ab = sum_i grad L_m(a,i) q_(ab,i)^T,  i = 0, 1
//           (q from the pair table of the reduced tables), so
//             G_ab = sum_i (B^T grad L_m(a,i)) (B^T q_(ab,i))^T
//           and the mass sum M_ab is a table entry too.  54 FMAs for G in
//           place of 108, no per-point loop.
// Lane roles: node row i of the wave, column node b of the element row,
// incidence part p (lanes per row = 10 S); the lane's b is fixed for the
// wave, so its reference gradients and mass values (KF = 0) live in
// registers; with KF = 1 the lane reads its pair row (a, b) of the table,
// 64 bytes, two steps ahead with the payload.
// In-row offset of a fixed local node b: the record word holding it is
// picked with per-lane masks (all-ones for that word) rather than an index,
// which keeps the record in registers.
template <int RB, int OFFW>
struct OffsetPick {
    static constexpr int NW = (RB - 4) / 4;
    uint32_t mask[NW];
    int shift;
    __device__ explicit OffsetPick(int b) {
        constexpr int per = 4 / OFFW;
#pragma unroll
        for (int t = 0; t < NW; ++t) mask[t] = (b / per == t) ? 0xffffffffu : 0u;
        shift = (b % per) * 8 * OFFW;
    }
    __device__ int operator()(const Record<RB>& r) const {
        uint32_t word = 0;
#pragma unroll
        for (int t = 0; t < NW; ++t) {
            uint32_t wt;
            __builtin_memcpy(&wt, r.off + 4 * t, 4);
            word |= wt & mask[t];
        }
        return int((word >> shift) & (OFFW == 1 ? 0xffu : 0xffffu));
    }
};

template <bool F32, int RB, int OFFW, int QG, int KF>
__global__ void __launch_bounds__(64) k_gather_vec10(GatherArgs A) {
    extern __shared__ double span_lds[];
    // two steps in flight: the PL_VEC payloads take 2 x 14 doubles per lane
    // (measured on 48^3 elasticity with KF = 0: K = 2 0.976 ms, K = 1 1.045 ms)
    // payload: 14 doubles, or 16 floats for the FP32-products form
    constexpr int NL = 10, C = 3, P = F32 ? 16 : 14, K = 2;
    constexpr int NQ = KF == 1 ? 8 : 1;   // pair-table values held per step
    using T = typename std::conditional<F32, float, double>::type;
    const int lane = threadIdx.x;
    const int wv = placed_wave(A.plan.xcd_seg, int(A.plan.nwaves), A.plan.placement, A.plan.xcd_chunk);
    if (wv < 0) return;
    const PlanArgs& PA = A.plan;
    const WaveMeta wm = PA.meta[wv];
    const int64_t r0w = wm.r0;
    Lane L;  // record addressing of load_records
    L.R = wm.R;
    L.S = wm.S;
    const int lpr = NL * L.S;
    L.i = lane / lpr;
    const int rem = lane - L.i * lpr;
    const int b = rem / L.S;
    L.p = rem - b * L.S;
    L.c = L.d = 0;
    L.lmul = L.S;
    L.ladd = L.p;
    L.active = L.i < L.R;
    L.rec0 = wm.rec0;
    L.steps = wm.steps;
    L.lo = wm.lo;
    L.span = int64_t(wm.span) * C * C;
    L.rb = 0;
    int row_len = 0;
    if (L.active) {
        const int64_t rn = r0w + L.i;
        const int64_t rbeg = PA.indptr[rn];
        row_len = int(PA.indptr[rn + 1] - rbeg);
        L.rb = int64_t(C) * C * (rbeg - L.lo);
    }
    const int64_t nsteps = (L.steps + L.S - 1) / L.S;
    const bool mass = A.has_mass != 0;
    // two buffer sets, used alternately (no register rotation): set X holds
    // the records, payloads and pair rows of the batch being computed, set Y
    // those of the next batch; a batch's records are reduced to what the
    // step needs (packed element/row, in-row offset) before their registers
    // take the records two batches ahead
    Record<RB> rX[K], rY[K];
    T pX[K][P], pY[K][P];
    T qX[K][NQ], qY[K][NQ];
    load_records<K, RB>(A, L, 0, rX);
    zero_span(span_lds, L.span, lane);
    const int Qm = A.Qm;
    const int lg = QG * NL * 3;
    const T* src = F32 ? reinterpret_cast<const T*>(A.tab32) : reinterpret_cast<const T*>(A.tab);
    // tables behind the span: KF = 0 the weighted rows w_q dphi_a(q) and
    // w_q phi_a(q) (lane-varying a); KF = 1 grad L_m(a,i) for the 20 (a, i)
    T* wdg = reinterpret_cast<T*>(span_lds + L.span);
    T* wpm = wdg + lg;
    if constexpr (KF == 1) {
        if (lane < 60) {
            constexpr int E0[6] = {0, 1, 2, 0, 1, 2}, E1[6] = {1, 2, 0, 3, 3, 3};
            const int a = lane / 6, i = (lane / 3) & 1, x = lane % 3;
            const int m = a < 4 ? a : (i == 0 ? E0[a - 4] : E1[a - 4]);
            wdg[lane] = m == 0 ? T(-1) : (x == m - 1 ? T(1) : T(0));
        }
    } else {
        for (int k = lane; k < lg + Qm * NL; k += 64) {
            if (k < lg)
                wdg[k] = src[QG + k] * src[k / (NL * 3)];
            else
                wpm[k - lg] = src[QG + lg + Qm + (k - lg)] * src[QG + lg + (k - lg) / NL];
        }
    }
    // KF = 0: the lane's column node b, unweighted gradients and values
    T gb[QG][3];
    if constexpr (KF == 0) {
#pragma unroll
        for (int q = 0; q < QG; ++q)
#pragma unroll
            for (int j = 0; j < 3; ++j) gb[q][j] = src[QG + (q * NL + b) * 3 + j];
    }
    // KF = 1: pair rows (a, b) of the table, 8 values each
    const T* qt = F32 ? reinterpret_cast<const T*>(A.qtab32) : reinterpret_cast<const T*>(A.qtab);
    auto load_pairs = [&](const Record<RB>* rec, T (*qv)[NQ]) {
        if constexpr (KF == 1) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (rec[k].packed < 0) continue;
                const T* row = qt + ((rec[k].packed & 15) * NL + b) * 8;
                if constexpr (F32) {
                    const float4 v0 = reinterpret_cast<const float4*>(row)[0];
                    const float4 v1 = reinterpret_cast<const float4*>(row)[1];
                    qv[k][0] = v0.x; qv[k][1] = v0.y; qv[k][2] = v0.z; qv[k][3] = v0.w;
                    qv[k][4] = v1.x; qv[k][5] = v1.y; qv[k][6] = v1.z; qv[k][7] = v1.w;
                } else {
                    const double2* r2v = reinterpret_cast<const double2*>(row);
#pragma unroll
                    for (int t = 0; t < 3; ++t) {
                        const double2 v = r2v[t];
                        qv[k][2 * t] = v.x;
                        qv[k][2 * t + 1] = v.y;
                    }
                    qv[k][6] = mass ? row[6] : 0.;
                    qv[k][7] = 0.;
                }
            }
        }
    };
    load_payloads<K, RB, P>(A, rX, pX);
    load_pairs(rX, qX);
    load_records<K, RB>(A, L, K, rY);
    __syncthreads();
    const T* phb = src + QG + lg + Qm + b;   // phi_b(q) at stride NL
    const OffsetPick<RB, OFFW> pick(b);
    struct StepRef {
        int32_t packed;   // element << 4 | local row a, -1 for padding
        int32_t off;      // in-row offset of the lane's column node b
    };
    auto reduce = [&](const Record<RB>* rec, StepRef* ex) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ex[k].packed = rec[k].packed;
            ex[k].off = pick(rec[k]);
        }
    };
    auto compute = [&](const StepRef* ex, const T (*pb)[P], const T (*qb)[NQ]) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (ex[k].packed < 0) continue;
            const T* pl = pb[k];
            const int a = ex[k].packed & 15;
            T B[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) B[t] = T(pl[t]);
            T G[3][3];
            T mab = T(0);
            if constexpr (KF == 1) {
                const T* pa = wdg + a * 6;
                const T* qa = qb[k];
                T u[2][3], w[2][3];
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        u[i][x] = fma(pa[3 * i], B[x], fma(pa[3 * i + 1], B[3 + x], pa[3 * i + 2] * B[6 + x]));
                        w[i][x] = fma(qa[3 * i], B[x], fma(qa[3 * i + 1], B[3 + x], qa[3 * i + 2] * B[6 + x]));
                    }
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int y = 0; y < 3; ++y) G[x][y] = fma(u[0][x], w[0][y], u[1][x] * w[1][y]);
                mab = qa[6];
            } else {
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int y = 0; y < 3; ++y) G[x][y] = T(0);
#pragma unroll
                for (int q = 0; q < QG; ++q) {
                    const T* ga = wdg + (q * NL + a) * 3;
                    const T g0 = ga[0], g1 = ga[1], g2 = ga[2];
                    T va[3], vb[3];
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        va[x] = fma(B[0 * 3 + x], g0, fma(B[1 * 3 + x], g1, B[2 * 3 + x] * g2));
                        vb[x] = fma(B[0 * 3 + x], gb[q][0], fma(B[1 * 3 + x], gb[q][1], B[2 * 3 + x] * gb[q][2]));
                    }
#pragma unroll
                    for (int x = 0; x < 3; ++x)
#pragma unroll
                        for (int y = 0; y < 3; ++y) G[x][y] = fma(va[x], vb[y], G[x][y]);
                }
                if (mass)
                    for (int q = 0; q < Qm; ++q) mab = fma(wpm[q * NL + a], phb[q * NL], mab);
            }
            const T iso = T(pl[9]), hs = T(pl[10]), lam = T(pl[11]);
            T diag = iso * ((G[0][0] + G[1][1]) + G[2][2]);
            if (mass) diag = fma(T(pl[12]), mab, diag);
            double v[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    T e = fma(lam, G[c][d], hs * G[d][c]);
                    if (c == d) e += diag;
                    v[c][d] = double(e);
                }
            double* col = span_lds + int(L.rb) + C * ex[k].off;
            auto add_block = [&]() {
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) lds_add(col + c * C * row_len + d, v[c][d]);
            };
            // the S parts of a row add in turn (one writer per address), or
            // at once with the LDS atomic serialising shared addresses
            if (L.S == 1 || PA.concurrent) {
                add_block();
            } else {
                for (int turn = 0; turn < L.S; ++turn)
                    if (L.p == turn) add_block();
            }
        }
    };
    for (int64_t st = 0; st < nsteps; st += 2 * K) {
        StepRef ex[K];
        // batch st from set X; set Y receives the payloads of batch st + K,
        // set X the records of batch st + 2K
        reduce(rX, ex);
        load_payloads<K, RB, P>(A, rY, pY);
        load_pairs(rY, qY);
        load_records<K, RB>(A, L, st + 2 * K, rX);
        compute(ex, pX, qX);
        if (st + K >= nsteps) break;
        // batch st + K from set Y; the roles swap
        reduce(rY, ex);
        load_payloads<K, RB, P>(A, rX, pX);
        load_pairs(rX, qX);
        load_records<K, RB>(A, L, st + 3 * K, rY);
        compute(ex, pY, qY);
    }
    __syncthreads();
    write_span(span_lds, L.span, lane, A.values + L.lo * C * C, A.plan.streaming);
}

// ------------------------------------------------------------------ MFMA gather
// Scalar fields (C == 1), element-constant coefficients, up to 4 points in the
// reduced gradient and mass rules.  A wave's rows are taken 16 at a time:
// lane l works on row slot l % 16 and quadrature point q = l / 16 and forms
//   u_q = D^T (w_q grad(phi_a)(q))          (the row's local node a)
// then one 16x16x4 MFMA per gradient component gives, for the 16 rows,
//   K^T[b][row] = sum_q grad(phi_b)(q)_x u_q(row)_x  (+ y, z, + mass)
// with the reference gradients as per-lane constants.  Lane l receives
// entries b = l/16 + 4r (FP64 layout) or 4(l/16) + r (FP32 layout) of its
// own row, so every entry is computed once and added by one lane: each LDS
// address is written by at most one lane per step and the per-address sum
// runs in the row's element order (deterministic).
typedef double mf_v4d __attribute__((ext_vector_type(4)));
typedef float mf_v4f __attribute__((ext_vector_type(4)));

template <bool F32>
struct MfmaT;
template <>
struct MfmaT<false> {
    using T = double;
    using V = mf_v4d;
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // entry b of output element r for lane group p
    static __device__ __forceinline__ int entry(int p, int r) { return p + 4 * r; }
};
template <>
struct MfmaT<true> {
    using T = float;
    using V = mf_v4f;
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int entry(int p, int r) { return 4 * p + r; }
};

template <int NL, int PL, bool F32, int RB, int OFFW>
__global__ void __launch_bounds__(64) k_gather_mfma(GatherArgs A) {
    extern __shared__ double span_lds[];
    using M = MfmaT<F32>;
    using T = typename M::T;
    constexpr int P = PayloadLayout<PL>::stride;
    constexpr int K = 2;   // (step, row group) units per pipeline batch
    const int lane = threadIdx.x;
    const int p = lane >> 4;      // quadrature point of this lane
    const int slot = lane & 15;   // row slot within a group of 16
    const int wv = placed_wave(A.plan.xcd_seg, int(A.plan.nwaves), A.plan.placement, A.plan.xcd_chunk);
    if (wv < 0) return;
    const int64_t r0w = A.plan.wave_r0[wv], r1w = A.plan.wave_r0[wv + 1];
    const int R = int(r1w - r0w);
    if (R < A.rmin || R > A.rmax) return;
    const int G = (R + 15) >> 4;
    const int64_t rec0 = A.plan.wave_rec[wv];
    const int64_t steps = (A.plan.wave_rec[wv + 1] - rec0) / R;
    const int64_t lo = A.plan.indptr[r0w];
    const int64_t span = A.plan.indptr[r1w] - lo;
    const int64_t units = steps * G;
    // tables: weighted rows for u and mass (lane-varying a), behind the span
    const int Qg = A.Qg, Qm = A.Qm;
    const T* src = F32 ? reinterpret_cast<const T*>(A.tab32) : reinterpret_cast<const T*>(A.tab);
    const int lg = Qg * NL * 3;
    T* wdg = reinterpret_cast<T*>(span_lds + ((span + 1) & ~int64_t(1)));
    T* wpm = wdg + lg;
    for (int k = lane; k < lg + Qm * NL; k += 64) {
        if (k < lg) wdg[k] = src[Qg + k] * src[k / (NL * 3)];
        else wpm[k - lg] = src[Qg + lg + Qm + (k - lg)] * src[Qg + lg + (k - lg) / NL];
    }
    zero_span(span_lds, span, lane);
    // per-lane MFMA A operands: grad(phi_b)(q) components and phi_b(q), b = slot, q = p
    T ax = T(0), ay = T(0), az = T(0), am = T(0);
    if (slot < NL && p < Qg) {
        const T* g = src + Qg + (p * NL + slot) * 3;
        ax = g[0];
        ay = g[1];
        az = g[2];
    }
    if (slot < NL && p < Qm) am = src[Qg + lg + Qm + p * NL + slot];
    const bool mass = A.has_mass != 0;
    // in-row offsets of the lane's entries b = entry(p, r), picked by masks
    const OffsetPick<RB, OFFW> pick[4] = {OffsetPick<RB, OFFW>(min(M::entry(p, 0), NL - 1)),
                                          OffsetPick<RB, OFFW>(min(M::entry(p, 1), NL - 1)),
                                          OffsetPick<RB, OFFW>(min(M::entry(p, 2), NL - 1)),
                                          OffsetPick<RB, OFFW>(min(M::entry(p, 3), NL - 1))};
    // unit t = (step t / G, group t % G): the lane's row is 16 (t % G) + slot
    auto load_unit = [&](int64_t t, Record<RB>& rec) {
        rec.packed = -1;
        if (t < units) {
            const int64_t st = t / G;
            const int row = int(t - st * G) * 16 + slot;
            if (row < R) load_record<RB>(A.plan.records + (rec0 + st * R + row) * RB, rec, A.plan.streaming);
        }
    };
    Record<RB> r0[K], r1[K], r2[K];
    double p0[K][P], p1[K][P];
#pragma unroll
    for (int k = 0; k < K; ++k) load_unit(k, r0[k]);
    load_payloads<K, RB, P>(A, r0, p0);
#pragma unroll
    for (int k = 0; k < K; ++k) load_unit(K + k, r1[k]);
    __syncthreads();
    for (int64_t t0 = 0; t0 < units; t0 += K) {
        load_payloads<K, RB, P>(A, r1, p1);
#pragma unroll
        for (int k = 0; k < K; ++k) load_unit(t0 + 2 * K + k, r2[k]);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int64_t t = t0 + k;
            if (t >= units) break;   // uniform
            const Record<RB>& rec = r0[k];
            const bool live = rec.packed >= 0;
            const int a = live ? (rec.packed & 15) : 0;
            const double* pl = p0[k];
            T D[9];
            T mf;
            if constexpr (PL == PL_FULL) {
#pragma unroll
                for (int u = 0; u < 9; ++u) D[u] = T(pl[u]);
                mf = T(pl[9]);
            } else {
                D[0] = T(pl[0]); D[1] = T(pl[1]); D[2] = T(pl[2]);
                D[3] = T(pl[1]); D[4] = T(pl[3]); D[5] = T(pl[4]);
                D[6] = T(pl[2]); D[7] = T(pl[4]); D[8] = T(pl[5]);
                mf = T(pl[6]);
            }
            T u0 = T(0), u1 = T(0), u2 = T(0), mb = T(0);
            if (live && p < Qg) {
                const T* ga = wdg + (p * NL + a) * 3;
                const T g0 = ga[0], g1 = ga[1], g2 = ga[2];
                u0 = fma(g0, D[0], fma(g1, D[3], g2 * D[6]));
                u1 = fma(g0, D[1], fma(g1, D[4], g2 * D[7]));
                u2 = fma(g0, D[2], fma(g1, D[5], g2 * D[8]));
            }
            if (mass && live && p < Qm) mb = mf * wpm[p * NL + a];
            typename M::V c = {T(0), T(0), T(0), T(0)};
            c = M::mma(ax, u0, c);
            c = M::mma(ay, u1, c);
            c = M::mma(az, u2, c);
            if (mass) c = M::mma(am, mb, c);
            if (live) {
                const int64_t st = t / G;
                const int row = int(t - st * G) * 16 + slot;
                const int64_t rb = A.plan.indptr[r0w + row] - lo;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int b = M::entry(p, r);
                    if (b < NL) lds_add(span_lds + rb + pick[r](rec), double(c[r]));
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            r0[k] = r1[k];
            r1[k] = r2[k];
#pragma unroll
            for (int u = 0; u < P; ++u) p0[k][u] = p1[k][u];
        }
    }
    __syncthreads();
    write_span(span_lds, span, lane, A.values + lo, A.plan.streaming);
}

// ------------------------------------------------------------------ vector-P2 MFMA gather
// 10-node 3-component fields (elasticity: lam div div + s eps:eps + iso
// grad:grad + m u.v), element-constant coefficients, 4-point reduced
// gradient rule, at most 4 mass points, on the 16-row wave layout (plan
// row_lanes = 4).  Lane l owns row slot n = l & 15 and contraction slice
// p = l >> 4.  With g = B^T (w_q grad(phi_a)(q)) for the row's element
// (payload B = adj/sqrt|det|) and local row a, every entry of the 3x3 node
// block (a, b) is a contraction over (q, k) with the reference gradients:
//   K_(a,c),(b,d) = sum_q sum_k grad(phi_b)(q)_k X_cd(q, k) + sum_q phi_b(q) X_cd(q, 3)
//   X_cd(q, k) = lam g_c B_kd + s/2 g_d B_kc + [c == d] iso (g . B_k)   (k < 3)
//   X_cd(q, 3) = [c == d] |det| m wm_q phi_a(q)
// so for each component pair cd and point q one 16x16x4 MFMA with the
// reference table as A (rows b, 10 of 16 used; k = p) and X as B (columns =
// the 16 rows of the wave) adds that point's share of all 10 blocks of all
// 16 rows: 36 MFMAs per incidence layer.  The accumulation over (q, k) runs
// in the MFMA's precision (FP32 or FP64); the lane then holds entries
// b = entry(p, r) of its own row and adds them to the FP64 span.  Each LDS
// address is written by one lane per instruction and the layers run in
// incidence order, so the per-address summation order is the row's element
// order.
template <bool F32, int RB, int OFFW>
__global__ void __launch_bounds__(64) k_gather_vec10_mfma(GatherArgs A) {
    extern __shared__ double span_lds[];
    using M = MfmaT<F32>;
    using T = typename M::T;
    constexpr int NL = 10, QG = 4, P = F32 ? 16 : 14;
    constexpr int lg = QG * NL * 3;
    const int lane = threadIdx.x;
    const int n = lane & 15, p = lane >> 4;
    const int wv = placed_wave(A.plan.xcd_seg, int(A.plan.nwaves), A.plan.placement, A.plan.xcd_chunk);
    if (wv < 0) return;
    const WaveMeta wm = A.plan.meta[wv];
    const int R = wm.R;
    const int64_t span = int64_t(wm.span) * 9;
    const int Qm = A.Qm;
    const T* src = F32 ? reinterpret_cast<const T*>(A.tab32) : reinterpret_cast<const T*>(A.tab);
    // weighted tables behind the span: w_q grad(phi_a)(q) and wm_q phi_a(q)
    T* wdg = reinterpret_cast<T*>(span_lds + span);
    T* wpm = wdg + lg;
    for (int k = lane; k < lg + Qm * NL; k += 64) {
        if (k < lg)
            wdg[k] = src[QG + k] * src[k / (NL * 3)];
        else
            wpm[k - lg] = src[QG + lg + Qm + (k - lg)] * src[QG + lg + (k - lg) / NL];
    }
    zero_span(span_lds, span, lane);
    // A operands: reference gradients (slices 0-2) and values (slice 3) of node b = n
    T aop[QG];
#pragma unroll
    for (int q = 0; q < QG; ++q) {
        aop[q] = T(0);
        if (n < NL) {
            if (p < 3) aop[q] = src[QG + (q * NL + n) * 3 + p];
            else if (q < Qm) aop[q] = src[QG + lg + Qm + q * NL + n];
        }
    }
    const bool live = n < R;
    int rlen = 0, rb = 0;
    if (live) {
        const int64_t b0 = A.plan.indptr[wm.r0 + n];
        rlen = int(A.plan.indptr[wm.r0 + n + 1] - b0);
        rb = int(9 * (b0 - wm.lo));
    }
    const int64_t steps = wm.steps;
    // in-row offsets of the lane's entries b = entry(p, r), picked by masks
    // (a runtime index into the record would put it in scratch memory)
    const OffsetPick<RB, OFFW> pick[4] = {OffsetPick<RB, OFFW>(min(M::entry(p, 0), NL - 1)),
                                          OffsetPick<RB, OFFW>(min(M::entry(p, 1), NL - 1)),
                                          OffsetPick<RB, OFFW>(min(M::entry(p, 2), NL - 1)),
                                          OffsetPick<RB, OFFW>(min(M::entry(p, 3), NL - 1))};
    auto load_rec = [&](int64_t layer, Record<RB>& r) {
        r.packed = -1;
        if (live && layer < steps)
            load_record<RB>(A.plan.records + (wm.rec0 + layer * R + n) * RB, r, A.plan.streaming);
    };
    Record<RB> r0[1], r1[1], r2[1];
    T p0[1][P], p1[1][P];
    load_rec(0, r0[0]);
    load_payloads<1, RB, P>(A, r0, p0);
    load_rec(1, r1[0]);
    __syncthreads();
    const T slice = p < 3 ? T(1) : T(0);
    for (int64_t st = 0; st < steps; ++st) {
        load_payloads<1, RB, P>(A, r1, p1);
        load_rec(st + 2, r2[0]);
        const bool ok = r0[0].packed >= 0;
        const int a = ok ? (r0[0].packed & 15) : 0;
        T B[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) B[t] = ok ? T(p0[0][t]) : T(0);
        const T iso = ok ? T(p0[0][9]) : T(0), hs = ok ? T(p0[0][10]) : T(0);
        const T lam = ok ? T(p0[0][11]) : T(0), mm = ok ? T(p0[0][12]) : T(0);
        // row p of B for the gradient slices; slice 3 carries the mass only
        T Bp[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) Bp[x] = slice * (p == 0 ? B[x] : p == 1 ? B[3 + x] : B[6 + x]);
        T lB[3], hB[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            lB[x] = lam * Bp[x];
            hB[x] = hs * Bp[x];
        }
        const T isop = slice * iso, mp = (T(1) - slice) * mm;
        typename M::V acc[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[t] = typename M::V{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int q = 0; q < QG; ++q) {
            const T* ga = wdg + (q * NL + a) * 3;
            const T h0 = ga[0], h1 = ga[1], h2 = ga[2];
            T g[3];
#pragma unroll
            for (int x = 0; x < 3; ++x) g[x] = fma(B[x], h0, fma(B[3 + x], h1, B[6 + x] * h2));
            const T gB = fma(g[0], Bp[0], fma(g[1], Bp[1], g[2] * Bp[2]));
            const T mt = q < Qm ? mp * wpm[q * NL + a] : T(0);
            const T dg = fma(isop, gB, mt);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    T X = fma(g[c], lB[d], g[d] * hB[c]);
                    if (c == d) X += dg;
                    acc[c * 3 + d] = M::mma(aop[q], X, acc[c * 3 + d]);
                }
        }
        if (ok) {
            int off[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) off[r] = pick[r](r0[0]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double* rowc = span_lds + rb + c * 3 * rlen;
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (M::entry(p, r) < NL) lds_add(rowc + 3 * off[r] + d, double(acc[c * 3 + d][r]));
            }
        }
        r0[0] = r1[0];
        r1[0] = r2[0];
#pragma unroll
        for (int t = 0; t < P; ++t) p0[0][t] = p1[0][t];
    }
    __syncthreads();
    write_span(span_lds, span, lane, A.values + wm.lo * 9, A.plan.streaming);
}

// ------------------------------------------------------------------ element-table gather
// Symmetric scalar path (C == 1, D symmetric): with affine elements and
// element-constant coefficients the element matrix is linear in the six
// entries of D and the mass factor,
//   K_ab = sum_k d_k N[a][b][k],  d = (D00, D01, D02, D11, D12, D22, |det| m),
//   N[a][b][k] = sum_q wg_q (dphi_ai dphi_bj + [i != j] dphi_aj dphi_bi)   (k = (i,j), i <= j)
//   N[a][b][6] = sum_q wm_q phi_a phi_b,
// so a row step costs 6-7 FMAs per entry.  The S lanes of a row divide the
// entries b = p, p + S, ... among themselves (no lane recomputes another's
// entries).  Within a step each LDS address is written by at most one lane
// (distinct b of one element are distinct columns), so the per-address
// summation order is the element order of the row's incidences.
// table slots per (a, b): 6 D entries and mass; the odd stride spreads the
// lanes' random (a, b) rows over the LDS banks
constexpr int kNK = 7;

__device__ inline void build_table_entry(const double* __restrict__ rtab, int n, int32_t Qg, int32_t Qm, int t,
                                         double* __restrict__ out, float* __restrict__ out32) {
    const int a = t / n, b = t % n;
    const double* wg = rtab;
    const double* dg = wg + Qg;
    const double* wm = dg + Qg * n * 3;
    const double* pm = wm + Qm;
    double M[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    for (int q = 0; q < Qg; ++q)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M[i][j] += wg[q] * dg[(q * n + a) * 3 + i] * dg[(q * n + b) * 3 + j];
    double mass = 0.;
    for (int q = 0; q < Qm; ++q) mass += wm[q] * pm[q * n + a] * pm[q * n + b];
    const double v[kNK] = {M[0][0], M[0][1] + M[1][0], M[0][2] + M[2][0], M[1][1], M[1][2] + M[2][1], M[2][2], mass};
    for (int k = 0; k < kNK; ++k) {
        out[t * kNK + k] = v[k];
        out32[t * kNK + k] = float(v[k]);
    }
}

template <int NL, bool F32, int RB, int OFFW, bool MASS>
__device__ inline void table_step(const GatherArgs& A, const Lane& L, const Record<RB>& rec, const double* pl,
                                  double* span_lds,
                                  const typename std::conditional<F32, float, double>::type* ntab) {
    using T = typename std::conditional<F32, float, double>::type;
    const int a = rec.packed & 15;
    const T d0 = T(pl[0]), d1 = T(pl[1]), d2 = T(pl[2]), d3 = T(pl[3]), d4 = T(pl[4]), d5 = T(pl[5]);
    const T d6 = MASS ? T(pl[6]) : T(0);
    const T* na = ntab + a * NL * kNK;
    for (int b = L.p; b < NL; b += L.S) {
        const T* n = na + b * kNK;
        T k = d0 * n[0];
        k = fma(d1, n[1], k);
        k = fma(d2, n[2], k);
        k = fma(d3, n[3], k);
        k = fma(d4, n[4], k);
        k = fma(d5, n[5], k);
        if (MASS) k = fma(d6, n[6], k);
        lds_add(span_lds + L.rb + rec_offset_dyn<RB, OFFW>(rec, b), double(k));
    }
}

// Workgroups of WPB waves share one copy of the element table; each wave
// owns a span region of `span_stride` doubles behind it.  Block placement
// follows the plan's XCD chunking at block granularity.
// LOAD: lane 0 of each row's lanes (it walks every incidence of the row, as
// all of them do) also sums |det J_e| w_a over the row's incidences, in
// incidence order, and writes b_i = f x that sum (the constant-source load
// vector of linear_form_assembler.cpp:41-137 on affine elements); |det J| is
// the payload's last slot.
template <int NL, bool F32, int RB, int OFFW, bool MASS, int WPB, bool LOAD = false>
__global__ void __launch_bounds__(64 * WPB) k_gather_table(GatherArgs A, int32_t span_stride, int32_t nblocks) {
    extern __shared__ double lds[];
    constexpr int P = PayloadLayout<PL_SYM>::stride;
    constexpr int K = 2;
    using T = typename std::conditional<F32, float, double>::type;
    constexpr int kTab = NL * NL * kNK;
    constexpr int kTabDoubles = ((kTab * int(sizeof(T)) + 15) / 16) * 2;
    T* ntab = reinterpret_cast<T*>(lds);
    const T* src = F32 ? reinterpret_cast<const T*>(A.tab32) : reinterpret_cast<const T*>(A.tab);
    for (int k = threadIdx.x; k < kTab; k += 64 * WPB) ntab[k] = src[k];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    // block placement (XCD chunks of blocks when the plan asks for chunking)
    int pb = blockIdx.x;
    if (A.plan.placement != 0 && A.plan.xcd_chunk > 0) {
        const int chunk = A.plan.xcd_chunk;
        const int x = pb & 7, k = pb >> 3;
        const int full = (nblocks / (8 * chunk)) * (8 * chunk);
        if (pb < full) pb = ((k / chunk) * 8 + x) * chunk + (k % chunk);
    }
    const int64_t wv = int64_t(pb) * WPB + wib;
    const bool live = wv < A.plan.nwaves;
    double* span_lds = lds + kTabDoubles + int64_t(wib) * span_stride;
    Lane L;
    Record<RB> r0[K], r1[K], r2[K];
    double p0[K][P], p1[K][P];
    if (live) {
        lane_setup<1>(A.plan, int(wv), lane, L);
        load_records<K, RB>(A, L, 0, r0);
        zero_span(span_lds, L.span, lane);
        load_payloads<K, RB, P>(A, r0, p0);
        load_records<K, RB>(A, L, K, r1);
    }
    __syncthreads();
    if (!live) return;
    // LOAD: w_a = sum_q W_q phi_a(X_q) (the order of k_linear_const)
    double lw0 = 0., lw1 = 0., lw2 = 0., lw3 = 0., lacc = 0.;
    if constexpr (LOAD) {
        static_assert(NL == 4, "the fused load vector is for 4-node elements");
        const Tables Tb = tables_at(A.load_tables, A.load_q, 4, 4);
        for (int q = 0; q < A.load_q; ++q) {
            lw0 += Tb.w[q] * Tb.phi[q * 4 + 0];
            lw1 += Tb.w[q] * Tb.phi[q * 4 + 1];
            lw2 += Tb.w[q] * Tb.phi[q * 4 + 2];
            lw3 += Tb.w[q] * Tb.phi[q * 4 + 3];
        }
    }
    for (int64_t st = 0; st < L.steps; st += K) {
        load_payloads<K, RB, P>(A, r1, p1);
        load_records<K, RB>(A, L, st + 2 * K, r2);
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (r0[k].packed >= 0) {
                table_step<NL, F32, RB, OFFW, MASS>(A, L, r0[k], p0[k], span_lds, ntab);
                if constexpr (LOAD) {
                    const int a = r0[k].packed & 15;
                    const double wa = a == 0 ? lw0 : a == 1 ? lw1 : a == 2 ? lw2 : lw3;
                    lacc += p0[k][7] * wa;
                }
            }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            r0[k] = r1[k];
            r1[k] = r2[k];
#pragma unroll
            for (int t = 0; t < P; ++t) p0[k][t] = p1[k][t];
        }
    }
    if constexpr (LOAD) {
        if (L.active && L.p == 0) {
            const int64_t row = A.plan.meta[wv].r0 + L.i;
            const int64_t dst = A.row_map ? int64_t(A.row_map[row]) : row;
            if (dst >= 0) A.load_out[dst] = lacc * A.load_f[0];
        }
        if (A.value_map) {
            write_span_mapped(span_lds, L.span, lane, A.value_map + L.lo, A.values, A.plan.streaming);
            return;
        }
    }
    // the span is this wave's own: its LDS operations complete in order
    if (A.plan.diag3)
        write_span_diag3(span_lds, A.plan.meta[wv], A.plan.indptr, lane, A.values, A.plan.streaming);
    else
        write_span(span_lds, L.span, lane, A.values + L.lo, A.plan.streaming);
}

// N1 rows are global edges; incidence (e, local edge a).
//   K_ab = s_a s_b [ sum_q w_q phi_aq^T Dm phi_bq + curl_a^T Dc curl_b ]
// with Dm = kappa adj adj^T/|det| and Dc = (sum_q w_q) mu J^T J/|det|: the
// covariant Piola map of hcurl_basis.py:355-364 integrated as in
// integrate_tet_n1 (edge_topology.cpp:196-231).
struct GatherN1Args {
    PlanArgs plan;
    const double* payload;
    // corner form (k_gather_n1_corners): geometry and element coefficients
    const double4* xyz;
    const int32_t* vconn;
    const double* kappa;
    const double* mu;
    const double* tab;  // w[Q] | phi[Q][6][3] | curl[6][3]
    int32_t Q;
    int32_t include_mass, include_curl;
    const double* kappa_qp;  // (E,Q) when the coefficients vary per point, else null
    const double* mu_qp;
    double* values;
};

// Element row a of one incidence (the lane's record and payload) into the
// wave's span.  QN = 4: the 4-point rule of element-constant coefficients at
// compile time; 0: A.Q points (per-point coefficients).
template <int RB, int OFFW, int QN>
__device__ inline void n1_step(const GatherN1Args& A, const Lane& L, const Record<RB>& rec, const double* pl,
                               double* span_lds, const double* phi_l, const double* curl_l) {
    const int Q = QN > 0 ? QN : A.Q;
    const double* W = A.tab;
    const double* phi = W + Q;
    const double* curl = phi + Q * 18;
    const int a = rec.packed & 15;
    const double* m6 = pl;
    const double* c6 = pl + 6;
    const int64_t bits = __double_as_longlong(pl[12]);
    double Kr[6];
#pragma unroll
    for (int b = 0; b < 6; ++b) Kr[b] = 0.;
    if (A.include_mass) {
        const int64_t e = rec.packed >> 4;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double* pa = phi_l + (q * 6 + a) * 3;
            const double f = A.kappa_qp ? W[q] * A.kappa_qp[e * Q + q] : W[q];
            const double u0 = f * (pa[0] * m6[0] + pa[1] * m6[1] + pa[2] * m6[2]);
            const double u1 = f * (pa[0] * m6[1] + pa[1] * m6[3] + pa[2] * m6[4]);
            const double u2 = f * (pa[0] * m6[2] + pa[1] * m6[4] + pa[2] * m6[5]);
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                const double* pb = phi + (q * 6 + b) * 3;
                Kr[b] += u0 * pb[0] + u1 * pb[1] + u2 * pb[2];
            }
        }
    }
    if (A.include_curl) {
        const double* ca = curl_l + a * 3;
        double cw = 1.;
        if (A.mu_qp) {
            const int64_t e = rec.packed >> 4;
            cw = 0.;
            for (int q = 0; q < Q; ++q) cw += W[q] * A.mu_qp[e * Q + q];
        }
        const double u0 = cw * (ca[0] * c6[0] + ca[1] * c6[1] + ca[2] * c6[2]);
        const double u1 = cw * (ca[0] * c6[1] + ca[1] * c6[3] + ca[2] * c6[4]);
        const double u2 = cw * (ca[0] * c6[2] + ca[1] * c6[4] + ca[2] * c6[5]);
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            const double* cb = curl + b * 3;
            Kr[b] += u0 * cb[0] + u1 * cb[1] + u2 * cb[2];
        }
    }
    const double sa = ((bits >> a) & 1) ? -1. : 1.;
#pragma unroll
    for (int b = 0; b < 6; ++b) Kr[b] *= sa * (((bits >> b) & 1) ? -1. : 1.);
    // one writer per address: the S parts of a row add in turn (or all at
    // once, the LDS atomic serialising shared addresses)
    double* row = span_lds + int(L.rb);
    if (L.S == 1 || A.plan.concurrent) {
#pragma unroll
        for (int b = 0; b < 6; ++b) lds_add(row + rec_offset<RB, OFFW>(rec, b), Kr[b]);
        return;
    }
    for (int turn = 0; turn < L.S; ++turn)
        if (L.p == turn) {
#pragma unroll
            for (int b = 0; b < 6; ++b) lds_add(row + rec_offset<RB, OFFW>(rec, b), Kr[b]);
        }
}

// Software-pipelined like k_gather_h1 (payloads one step ahead, records
// two), with the incidence split: part p of a row takes its incidences
// p, p+S, ... and the parts add in turn, so every CSR entry receives its
// element contributions in incidence order.
template <int RB, int OFFW, int QN>
__global__ void __launch_bounds__(64) k_gather_n1(GatherN1Args A) {
    extern __shared__ double span_lds[];
    constexpr int K = 1, P = 14;   // a 14-double payload per lane and step
    const int lane = threadIdx.x;
    Lane L;
    const int wv = placed_wave(A.plan.xcd_seg, int(A.plan.nwaves), A.plan.placement, A.plan.xcd_chunk);
    if (wv < 0) return;
    lane_setup<1>(A.plan, wv, lane, L);
    L.lmul = L.S;
    L.ladd = L.p;
    const int64_t nsteps = (L.steps + L.S - 1) / L.S;
    Record<RB> r0[K], r1[K], r2[K];
    double p0[K][P], p1[K][P];
    load_records<K, RB>(A, L, 0, r0);
    zero_span(span_lds, L.span, lane);
    const int Q = QN > 0 ? QN : A.Q;
    const int tab_len = Q + Q * 18 + 18;
    double* tab_lds = span_lds + L.span;
    for (int k = lane; k < tab_len; k += 64) tab_lds[k] = A.tab[k];
    load_payloads<K, RB, P>(A, r0, p0);
    load_records<K, RB>(A, L, K, r1);
    __syncthreads();
    const double* phi_l = tab_lds + Q;
    const double* curl_l = phi_l + Q * 18;
    for (int64_t st = 0; st < nsteps; st += K) {
        load_payloads<K, RB, P>(A, r1, p1);
        load_records<K, RB>(A, L, st + 2 * K, r2);
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (r0[k].packed >= 0) n1_step<RB, OFFW, QN>(A, L, r0[k], p0[k], span_lds, phi_l, curl_l);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            r0[k] = r1[k];
            r1[k] = r2[k];
#pragma unroll
            for (int t = 0; t < P; ++t) p0[k][t] = p1[k][t];
        }
    }
    __syncthreads();
    write_span(span_lds, L.span, lane, A.values + L.lo, A.plan.streaming);
}


// Corner form of the N1 gather (element-constant coefficients, 4-point
// rule): no payload pass.  Each step forms its element's payload from the
// element's four corners (J, adjugate, kappa adj adj^T/|det|, mu wsum
// J^T J/|det| and the edge signs from the vertex ids, exactly as
// k_payload_n1 does) and then runs n1_step, so the values are bit-identical
// to the payload form.  Records are loaded two steps ahead and the corner ids
// and coefficients one step ahead; a step waits only for its corner
// coordinates, which neighbouring rows keep in L2.  C4 96^3: 0.693 ms with
// the payload pass (132 us) against 0.670 ms.
template <int RB, int OFFW>
__global__ void __launch_bounds__(64) k_gather_n1_corners(GatherN1Args A) {
    extern __shared__ double span_lds[];
    constexpr int QN = 4;
    const int lane = threadIdx.x;
    Lane L;
    const int wv = placed_wave(A.plan.xcd_seg, int(A.plan.nwaves), A.plan.placement, A.plan.xcd_chunk);
    if (wv < 0) return;
    lane_setup<1>(A.plan, wv, lane, L);
    L.lmul = L.S;
    L.ladd = L.p;
    const int64_t nsteps = (L.steps + L.S - 1) / L.S;
    Record<RB> r0[1], r1[1], r2[1];
    int4 c0 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
    double k0 = 0., k1 = 0., m0 = 0., m1 = 0.;
    load_records<1, RB>(A, L, 0, r0);
    zero_span(span_lds, L.span, lane);
    const int tab_len = QN + QN * 18 + 18;
    double* tab_lds = span_lds + L.span;
    for (int k = lane; k < tab_len; k += 64) tab_lds[k] = A.tab[k];
    double wsum = 0.;
#pragma unroll
    for (int q = 0; q < QN; ++q) wsum += A.tab[q];
    auto load_corners = [&](const Record<RB>& r, int4& c, double& kap, double& mu) {
        if (r.packed >= 0) {
            const int64_t e = r.packed >> 4;
            c = reinterpret_cast<const int4*>(A.vconn)[e];
            kap = A.include_mass ? A.kappa[e] : 0.;
            mu = A.include_curl ? A.mu[e] * wsum : 0.;
        }
    };
    load_corners(r0[0], c0, k0, m0);
    load_records<1, RB>(A, L, 1, r1);
    __syncthreads();
    const double* phi_l = tab_lds + QN;
    const double* curl_l = phi_l + QN * 18;
    for (int64_t st = 0; st < nsteps; ++st) {
        load_corners(r1[0], c1, k1, m1);
        load_records<1, RB>(A, L, st + 2, r2);
        if (r0[0].packed >= 0) {
            double J[9];
            const int32_t v[4] = {c0.x, c0.y, c0.z, c0.w};
            affine_jacobian(A.xyz, v, J);
            Geo G;
            adjugate(J, G);
            const double inv = 1. / fabs(G.det);
            const double* a = G.adj;
            double pl[14];
            const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                const int i = ij[t][0], j = ij[t][1];
                pl[t] = k0 * inv * (a[i * 3 + 0] * a[j * 3 + 0] + a[i * 3 + 1] * a[j * 3 + 1] + a[i * 3 + 2] * a[j * 3 + 2]);
                pl[6 + t] = m0 * inv * (J[0 * 3 + i] * J[0 * 3 + j] + J[1 * 3 + i] * J[1 * 3 + j] + J[2 * 3 + i] * J[2 * 3 + j]);
            }
            int64_t bits = 0;
#pragma unroll
            for (int l = 0; l < 6; ++l) bits |= int64_t(v[kN1Edge[l][0]] < v[kN1Edge[l][1]] ? 0 : 1) << l;
            pl[12] = __longlong_as_double(bits);
            pl[13] = 0.;
            n1_step<RB, OFFW, QN>(A, L, r0[0], pl, span_lds, phi_l, curl_l);
        }
        r0[0] = r1[0];
        r1[0] = r2[0];
        c0 = c1;
        k0 = k1;
        m0 = m1;
    }
    __syncthreads();
    write_span(span_lds, L.span, lane, A.values + L.lo, A.plan.streaming);
}

// ------------------------------------------------------------------ general H1 kernel
// Per-quadrature-point coefficients and/or isoparametric (curved P2)
// geometry: J, det and the coefficient are evaluated per point inside the
// gather step (basis_geometry.cpp:107-176 fused with the element kernel).
struct H1Args {
    PlanArgs plan;
    const double4* xyz;
    const int32_t* vconn;
    const int32_t* gconn;
    int32_t gn;
    const double* tables;  // w | phi | dphi | gphi | gdphi
    int32_t Q;
    const double *cm, *cg, *cs, *cl, *cA;
    int32_t per_qp;
    double* values;
};

__device__ inline void load_coef(const H1Args& A, int64_t idx, Coef& k) {
    k.m = A.cm ? A.cm[idx] : 0.;
    k.g = A.cg ? A.cg[idx] : 0.;
    k.s = A.cs ? A.cs[idx] : 0.;
    k.lam = A.cl ? A.cl[idx] : 0.;
    if (A.cA == nullptr) {
#pragma unroll
        for (int t = 0; t < 9; ++t) k.A[t] = 0.;
        return;
    }
    // 72-byte rows: 16-byte aligned at even idx (uniform over a wave for
    // per-point rows of an even Q), so four 16-byte loads and one 8-byte load
    const double* src = A.cA + idx * 9;
    if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
        const double2* v = reinterpret_cast<const double2*>(src);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double2 x = v[t];
            k.A[2 * t] = x.x;
            k.A[2 * t + 1] = x.y;
        }
        k.A[8] = src[8];
    } else {
        k.A[0] = src[0];
        const double2* v = reinterpret_cast<const double2*>(src + 1);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double2 x = v[t];
            k.A[2 * t + 1] = x.x;
            k.A[2 * t + 2] = x.y;
        }
    }
}

template <int C>
__device__ inline void build_D(const Geo& G, const Coef& k, int c, int d, bool has_tensor, bool has_scalar,
                               double* D) {
    const double inv = 1. / fabs(G.det);
    const double* a = G.adj;
    if (!has_scalar) {
        // tensor-only forms: no gradient, symmetric-gradient or divergence terms
#pragma unroll
        for (int t = 0; t < 9; ++t) D[t] = 0.;
    } else {
    const double iso = (c == d) ? (k.g + 0.5 * k.s) : 0.;
    // columns c and d of the adjugate: compile-time for one component,
    // else exact 0/1 weights (a runtime register index would go to scratch)
    double ac[3], ad[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if constexpr (C == 1) {
            ac[i] = ad[i] = a[i * 3];
        } else {
            ac[i] = a[i * 3] * double(c == 0) + a[i * 3 + 1] * double(c == 1) + a[i * 3 + 2] * double(c == 2);
            ad[i] = a[i * 3] * double(d == 0) + a[i * 3 + 1] * double(d == 1) + a[i * 3 + 2] * double(d == 2);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double aat = a[i * 3 + 0] * a[j * 3 + 0] + a[i * 3 + 1] * a[j * 3 + 1] + a[i * 3 + 2] * a[j * 3 + 2];
            D[i * 3 + j] = iso * aat + 0.5 * k.s * ad[i] * ac[j] + k.lam * ac[i] * ad[j];
        }
    }
    if (C == 1 && has_tensor) {
        double T[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                T[i * 3 + j] = a[i * 3 + 0] * k.A[0 * 3 + j] + a[i * 3 + 1] * k.A[1 * 3 + j] + a[i * 3 + 2] * k.A[2 * 3 + j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                D[i * 3 + j] += T[i * 3 + 0] * a[j * 3 + 0] + T[i * 3 + 1] * a[j * 3 + 1] + T[i * 3 + 2] * a[j * 3 + 2];
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) D[t] *= inv;
}

template <int NL, int C, bool ISO, int RB, int OFFW>
__global__ void __launch_bounds__(64) k_assemble_h1_general(H1Args A) {
    extern __shared__ double span_lds[];
    const int lane = threadIdx.x;
    Lane L;
    const int wv = placed_wave(A.plan.xcd_seg, int(A.plan.nwaves), A.plan.placement, A.plan.xcd_chunk);
    if (wv < 0) return;
    lane_setup<C>(A.plan, wv, lane, L);
    zero_span(span_lds, L.span, lane);
    __syncthreads();
    const int Q = A.Q;
    const Tables T = tables_at(A.tables, Q, NL, A.gn);
    const bool has_tensor = A.cA != nullptr;
    const bool has_mass = A.cm != nullptr;
    const bool has_scalar = A.cg != nullptr || A.cs != nullptr || A.cl != nullptr;
    // incidence split: part p of a row takes its incidence layers p, p+S, ...
    // (the plan pads the layers to a multiple of S) and the S parts add their
    // element rows in turn, so every entry sums in the row's element order
    const int64_t nsteps = (L.steps + L.S - 1) / L.S;
    for (int64_t st = 0; st < nsteps; ++st) {
        const int64_t layer = st * L.S + L.p;
        Record<RB> rec;
        rec.packed = -1;
        if (L.active && layer < L.steps) load_record<RB>(A.plan.records + (L.rec0 + layer * L.R + L.i) * RB, rec);
        const bool valid = rec.packed >= 0;
        double Kr[NL];
#pragma unroll
        for (int b = 0; b < NL; ++b) Kr[b] = 0.;
        if (valid) {
        const int64_t e = rec.packed >> 4;
        const int a = rec.packed & 15;
        Geo G;
        double Jm[9];
        Coef k;
        double D[9];
        double mfac = 0.;
        if (!ISO) {
            affine_jacobian(A.xyz, A.vconn + e * 4, Jm);
            adjugate(Jm, G);
        }
        for (int q = 0; q < Q; ++q) {
            if (ISO) {
                iso_jacobian(A.xyz, A.gconn + e * A.gn, A.gn, T.gdphi + q * A.gn * 3, Jm);
                adjugate(Jm, G);
            }
            load_coef(A, A.per_qp ? e * Q + q : e, k);
            build_D<C>(G, k, L.c, L.d, has_tensor, has_scalar, D);
            mfac = (L.c == L.d) ? fabs(G.det) * k.m : 0.;
            const double wq = T.w[q];
            const double* ga = T.dphi + (q * NL + a) * 3;
            const double u0 = wq * (ga[0] * D[0] + ga[1] * D[3] + ga[2] * D[6]);
            const double u1 = wq * (ga[0] * D[1] + ga[1] * D[4] + ga[2] * D[7]);
            const double u2 = wq * (ga[0] * D[2] + ga[1] * D[5] + ga[2] * D[8]);
#pragma unroll
            for (int b = 0; b < NL; ++b) {
                const double* gb = T.dphi + (q * NL + b) * 3;
                Kr[b] += u0 * gb[0] + u1 * gb[1] + u2 * gb[2];
            }
            if (has_mass) {
                const double v = wq * mfac * T.phi[q * NL + a];
#pragma unroll
                for (int b = 0; b < NL; ++b) Kr[b] += v * T.phi[q * NL + b];
            }
        }
        }
        for (int turn = 0; turn < L.S; ++turn) {
            if (valid && L.p == turn) {
#pragma unroll
                for (int b = 0; b < NL; ++b) {
                    const int o = rec_offset<RB, OFFW>(rec, b);
                    lds_add(span_lds + L.rb + int64_t(C) * o, Kr[b]);
                }
            }
        }
    }
    __syncthreads();
    write_span(span_lds, L.span, lane, A.values + L.lo * C * C, A.plan.streaming);
}

// ------------------------------------------------------------------ linear form
// out[node*C + c] = sum over incidences (e,a) of the node, over q, of
//   w_q |det_q| (phi_aq f_c + grad(phi_a) . G_c)
// f: constant (C) or per (e,q,c); G: constant (C,3) or per (e,q,c,3).
struct LinArgs {
    const double4* xyz;
    const int32_t* vconn;
    const int32_t* gconn;
    int32_t gn;
    const double* tables;
    int32_t Q, n, C;
    const double* f;
    int32_t f_per_qp;
    const double* G;
    int32_t G_per_qp;
    const int64_t* inc_ptr;
    const int32_t* inc;
    int64_t nnodes;
    double* out;
};

template <bool ISO>
__global__ void k_assemble_linear(LinArgs A) {
    const int64_t total = A.nnodes * A.C;
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < total; r += int64_t(gridDim.x) * blockDim.x) {
        const int64_t node = r / A.C;
        const int c = int(r - node * A.C);
        const Tables T = tables_at(A.tables, A.Q, A.n, A.gn);
        double acc = 0.;
        for (int64_t k = A.inc_ptr[node]; k < A.inc_ptr[node + 1]; ++k) {
            const int32_t packed = A.inc[k];
            const int64_t e = packed >> 4;
            const int a = packed & 15;
            double Jm[9];
            Geo G;
            if (!ISO) {
                affine_jacobian(A.xyz, A.vconn + e * 4, Jm);
                adjugate(Jm, G);
            }
            double local = 0.;
            for (int q = 0; q < A.Q; ++q) {
                if (ISO) {
                    iso_jacobian(A.xyz, A.gconn + e * A.gn, A.gn, T.gdphi + q * A.gn * 3, Jm);
                    adjugate(Jm, G);
                }
                double con = 0.;
                if (A.f) con += T.phi[q * A.n + a] * (A.f_per_qp ? A.f[(e * A.Q + q) * A.C + c] : A.f[c]);
                if (A.G) {
                    const double* Gc = A.G_per_qp ? A.G + ((e * A.Q + q) * A.C + c) * 3 : A.G + c * 3;
                    const double* ga = T.dphi + (q * A.n + a) * 3;
                    // grad(phi_a)_j = sum_i adj[i][j] ga_i / det
                    for (int j = 0; j < 3; ++j) {
                        const double gj = (G.adj[0 * 3 + j] * ga[0] + G.adj[1 * 3 + j] * ga[1] + G.adj[2 * 3 + j] * ga[2]) / G.det;
                        con += gj * Gc[j];
                    }
                }
                local += T.w[q] * fabs(G.det) * con;
            }
            acc += local;
        }
        A.out[r] = acc;
    }
}

// Constant load on affine elements: out[node*C + c] = f_c sum over the
// node's incidences (e,a) of |det J_e| sum_q w_q phi_a(q).
__global__ void k_elem_absdet(const double4* __restrict__ xyz, const int32_t* __restrict__ vconn, int64_t E,
                              double* __restrict__ out) {
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E; e += int64_t(gridDim.x) * blockDim.x) {
        double Jm[9];
        Geo G;
        affine_jacobian(xyz, vconn + e * 4, Jm);
        adjugate(Jm, G);
        out[e] = fabs(G.det);
    }
}

__global__ void k_linear_const(const int64_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc, int64_t nnodes,
                               int32_t C, const double* __restrict__ det, const double* __restrict__ tables,
                               int32_t Q, int32_t n, int32_t gn, const double* __restrict__ f,
                               double* __restrict__ out) {
    __shared__ double wphi[16];
    if (threadIdx.x < n) {
        const Tables T = tables_at(tables, Q, n, gn);
        double s = 0.;
        for (int q = 0; q < Q; ++q) s += T.w[q] * T.phi[q * n + threadIdx.x];
        wphi[threadIdx.x] = s;
    }
    __syncthreads();
    // 16 lanes per node: lane l sums incidences l, l+16, ..., then a fixed
    // shuffle tree adds the 16 partial sums (deterministic order); the loop
    // bound is uniform over each 16-lane group
    const int sub = threadIdx.x & 15;
    const int64_t groups = (int64_t(gridDim.x) * blockDim.x) >> 4;
    for (int64_t node = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4; node < nnodes; node += groups) {
        double acc = 0.;
        for (int64_t k = inc_ptr[node] + sub; k < inc_ptr[node + 1]; k += 16) {
            const int32_t packed = inc[k];
            acc += det[packed >> 4] * wphi[packed & 15];
        }
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        acc += __shfl_xor(acc, 4);
        acc += __shfl_xor(acc, 8);
        if (sub < C) out[node * C + sub] = acc * f[sub];
    }
}

// ------------------------------------------------------------------ geometry
// Per (e,q): det, tolerance 64*eps*max|J|^3, condition number, dx, physical
// point, and optionally the tabulated shape/gradients
// (basis_geometry.cpp:107-176).
struct GeoArgs {
    const double4* xyz;
    const int32_t* vconn;
    const int32_t* gconn;
    int32_t gn;
    const double* tables;
    int32_t Q, n;
    int64_t E;
    double *det, *tol, *cond, *dx, *x, *shape, *grads;
};

template <bool ISO>
__global__ void k_geometry(GeoArgs A) {
    const int64_t total = A.E * A.Q;
    const Tables T = tables_at(A.tables, A.Q, A.n, A.gn);
    for (int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; it < total; it += int64_t(gridDim.x) * blockDim.x) {
        const int64_t e = it / A.Q;
        const int q = int(it - e * A.Q);
        double J[9];
        if (ISO)
            iso_jacobian(A.xyz, A.gconn + e * A.gn, A.gn, T.gdphi + q * A.gn * 3, J);
        else
            affine_jacobian(A.xyz, A.vconn + e * 4, J);
        if (A.x) {
            double x0 = 0., x1 = 0., x2 = 0.;
            const int32_t* cg = ISO ? A.gconn + e * A.gn : A.vconn + e * 4;
            const int gn = ISO ? A.gn : 4;
            for (int a2 = 0; a2 < gn; ++a2) {
                const double4 p = A.xyz[cg[a2]];
                const double s = T.gphi[q * A.gn + a2];
                x0 += s * p.x; x1 += s * p.y; x2 += s * p.z;
            }
            A.x[it * 3 + 0] = x0; A.x[it * 3 + 1] = x1; A.x[it * 3 + 2] = x2;
        }
        double sc = 0.;
#pragma unroll
        for (int k = 0; k < 9; ++k) sc = fmax(sc, fabs(J[k]));
        const double tol = 64. * 2.220446049250313e-16 * sc * sc * sc;
        Geo G;
        adjugate(J, G);
        const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                           J[2] * (J[3] * J[7] - J[4] * J[6]);
        if (A.det) A.det[it] = det;
        if (A.tol) A.tol[it] = tol;
        const bool ok = isfinite(det) && isfinite(tol) && fabs(det) > tol;
        double jn = 0., in = 0.;
        double Ji[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            Ji[k] = G.adj[k] / det;
            jn += J[k] * J[k];
            in += Ji[k] * Ji[k];
        }
        if (A.cond) A.cond[it] = ok ? sqrt(jn * in) : 0.;
        if (A.dx) A.dx[it] = ok ? fabs(det) * T.w[q] : 0.;
        if (A.shape)
            for (int a2 = 0; a2 < A.n; ++a2) A.shape[it * A.n + a2] = ok ? T.phi[q * A.n + a2] : 0.;
        if (A.grads)
            for (int a2 = 0; a2 < A.n; ++a2)
                for (int j = 0; j < 3; ++j) {
                    const double* g = T.dphi + (q * A.n + a2) * 3;
                    A.grads[(it * A.n + a2) * 3 + j] = ok ? g[0] * Ji[0 * 3 + j] + g[1] * Ji[1 * 3 + j] + g[2] * Ji[2 * 3 + j] : 0.;
                }
    }
}

// ------------------------------------------------------------------ dispatch
static PlanArgs plan_args(const skfa_plan* P) {
    PlanArgs A;
    A.indptr = P->indptr.p;
    A.wave_r0 = P->wave_r0.p;
    A.wave_rec = P->wave_rec.p;
    A.wave_split = P->wave_split.p;
    A.records = P->records.p;
    A.nnodes = P->nnodes;
    A.nwaves = P->nwaves;
    A.xcd_seg = P->xcd_seg.p;
    A.meta = P->wave_meta.p;
    A.placement = P->tune_placement;
    A.xcd_chunk = P->tune_xcd_chunk;
    A.streaming = P->tune_streaming;
    A.concurrent = P->tune_split == SP_CONCURRENT;
    return A;
}

static dim3 wave_grid(const skfa_plan* P) {
    if (P->tune_placement == 1) {
        int64_t mx = 0;
        for (int x = 0; x < 8; ++x) mx = std::max<int64_t>(mx, P->xcd_seg_host[x + 1] - P->xcd_seg_host[x]);
        return dim3(unsigned(8 * mx));
    }
    return dim3(unsigned(P->nwaves));
}

static unsigned elem_grid(int64_t E) { return unsigned(std::min<int64_t>((E + 255) / 256, 65536)); }

static int ensure_payload(skfa_plan* P, int stride) {
    const size_t need = size_t(P->E) * stride;
    if (P->payload.n >= need && need) return kOk;
    return P->payload.alloc(need ? need : 1);
}

// Record layouts a plan produces (4 + n * offset bytes, rounded up to 8, 16
// or 32): 4-node elements (8, 1) and (16, 2), 10-node elements (16, 1) and
// (32, 2).  Only those are instantiated.
template <int NL, typename F>
static int with_record_layout(const skfa_plan* P, F&& launch) {
    using std::integral_constant;
    if constexpr (NL == 4) {
        if (P->rec_bytes == 8 && P->off_bytes == 1) return launch(integral_constant<int, 8>{}, integral_constant<int, 1>{});
        if (P->rec_bytes == 16 && P->off_bytes == 2) return launch(integral_constant<int, 16>{}, integral_constant<int, 2>{});
    } else {
        if (P->rec_bytes == 16 && P->off_bytes == 1) return launch(integral_constant<int, 16>{}, integral_constant<int, 1>{});
        if (P->rec_bytes == 32 && P->off_bytes == 2) return launch(integral_constant<int, 32>{}, integral_constant<int, 2>{});
    }
    set_error("unsupported gather record layout");
    return kUnsupported;
}

template <int NL, int C, int PL, bool F32, int QG, int SP, int KF = 0>
static int launch_gather_k(const skfa_plan* P, const GatherArgs& A, size_t lds, hipStream_t s) {
    const dim3 grid = wave_grid(P), block(64);
    if (SP == SP_ENTRY) lds += (size_t(A.Qg) * NL * 3 + size_t(A.Qm) * NL) * (F32 ? 4 : 8);
    return with_record_layout<NL>(P, [&](auto rb, auto ow) {
        constexpr int RB = decltype(rb)::value, OFFW = decltype(ow)::value;
        k_gather_h1<NL, C, PL, F32, RB, OFFW, QG, SP, KF><<<grid, block, lds, s>>>(A);
        SKFA_LAUNCH_CHECK();
        return kOk;
    });
}

// structural P2 rows (tuning key 8): scalar 10-node fields with the 4-point
// reduced gradient rule, when the host supplied the coefficient tables
template <int NL, int C, int QG>
static bool structural_rows(const skfa_plan* P, const GatherArgs& A) {
    return NL == 10 && C == 1 && QG == 4 && A.ctab != nullptr && P->tune_kform != 0;
}

// how the S lanes of a row share its work, from the plan's tuning
template <int NL, int C, int PL, bool F32, int QG>
static int launch_gather_q(const skfa_plan* P, const GatherArgs& A, size_t lds, hipStream_t s) {
    if (P->tune_split == SP_ENTRY) return launch_gather_k<NL, C, PL, F32, QG, SP_ENTRY>(P, A, lds, s);
    if (P->tune_split == SP_COLUMN) return launch_gather_k<NL, C, PL, F32, QG, SP_COLUMN>(P, A, lds, s);
    const bool conc = P->tune_split == SP_CONCURRENT;
    if constexpr (NL == 10 && C == 1 && QG == 4 && (PL == PL_SYM6 || PL == PL_SYM || PL == PL_SYM32)) {
        if (structural_rows<NL, C, QG>(P, A))
            return conc ? launch_gather_k<NL, C, PL, F32, QG, SP_CONCURRENT, 1>(P, A, lds, s)
                        : launch_gather_k<NL, C, PL, F32, QG, SP_INCIDENCE, 1>(P, A, lds, s);
    }
    return conc ? launch_gather_k<NL, C, PL, F32, QG, SP_CONCURRENT>(P, A, lds, s)
                : launch_gather_k<NL, C, PL, F32, QG, SP_INCIDENCE>(P, A, lds, s);
}

template <int NL, int C, int PL, bool F32>
static int launch_gather_pl(const skfa_plan* P, const GatherArgs& A, size_t lds, hipStream_t s) {
    if (A.Qg == 1) return launch_gather_q<NL, C, PL, F32, 1>(P, A, lds, s);
    if (A.Qg == 4) return launch_gather_q<NL, C, PL, F32, 4>(P, A, lds, s);
    return launch_gather_q<NL, C, PL, F32, 0>(P, A, lds, s);
}

constexpr int kTableWaves = 4;

template <int NL, bool F32, bool MASS, bool LOAD = false>
static int launch_table(const skfa_plan* P, const GatherArgs& A, hipStream_t s) {
    using T = typename std::conditional<F32, float, double>::type;
    constexpr int W = kTableWaves;
    const int32_t span_stride = int32_t(((P->lds_bytes / 8) + 1) & ~1);
    const int32_t nblocks = int32_t((P->nwaves + W - 1) / W);
    const size_t tab_bytes = size_t((NL * NL * kNK * sizeof(T) + 15) / 16) * 16;
    const size_t lds = tab_bytes + size_t(W) * span_stride * 8;
    SKFA_REQUIRE(lds <= 160 * 1024, "LDS request exceeds 160 KiB");
    const dim3 grid(unsigned(std::max(nblocks, 1))), block(64 * W);
    return with_record_layout<NL>(P, [&](auto rb, auto ow) {
        constexpr int RB = decltype(rb)::value, OFFW = decltype(ow)::value;
        k_gather_table<NL, F32, RB, OFFW, MASS, W, LOAD><<<grid, block, lds, s>>>(A, span_stride, nblocks);
        SKFA_LAUNCH_CHECK();
        return kOk;
    });
}

template <int NL, int PL, bool F32>
static int launch_mfma(const skfa_plan* P, const GatherArgs& A, hipStream_t s) {
    const dim3 grid = wave_grid(P), block(64);
    const size_t span = size_t((P->lds_bytes / 8 + 1) & ~1) * 8;
    const size_t lds = span + (size_t(A.Qg) * NL * 3 + size_t(A.Qm) * NL) * (F32 ? 4 : 8);
    SKFA_REQUIRE(lds <= 160 * 1024, "LDS request exceeds 160 KiB");
    return with_record_layout<NL>(P, [&](auto rb, auto ow) {
        constexpr int RB = decltype(rb)::value, OFFW = decltype(ow)::value;
        k_gather_mfma<NL, PL, F32, RB, OFFW><<<grid, block, lds, s>>>(A);
        SKFA_LAUNCH_CHECK();
        return kOk;
    });
}

template <bool F32>
static int launch_vec10(const skfa_plan* P, const GatherArgs& A, size_t lds, hipStream_t s) {
    const dim3 grid = wave_grid(P), block(64);
    // barycentric pair form when the pair table came with the reduced
    // tables (row form -1 or 1)
    const bool pairs = A.qtab != nullptr && (!F32 || A.qtab32 != nullptr) && P->tune_kform != 0;
    return with_record_layout<10>(P, [&](auto rb, auto ow) {
        constexpr int RB = decltype(rb)::value, OFFW = decltype(ow)::value;
        if (pairs)
            k_gather_vec10<F32, RB, OFFW, 4, 1><<<grid, block, lds, s>>>(A);
        else
            k_gather_vec10<F32, RB, OFFW, 4, 0><<<grid, block, lds, s>>>(A);
        SKFA_LAUNCH_CHECK();
        return kOk;
    });
}

template <bool F32>
static int launch_vec10_mfma(const skfa_plan* P, const GatherArgs& A, size_t lds, hipStream_t s) {
    const dim3 grid = wave_grid(P), block(64);
    return with_record_layout<10>(P, [&](auto rb, auto ow) {
        constexpr int RB = decltype(rb)::value, OFFW = decltype(ow)::value;
        k_gather_vec10_mfma<F32, RB, OFFW><<<grid, block, lds, s>>>(A);
        SKFA_LAUNCH_CHECK();
        return kOk;
    });
}

// the 16-row vector-P2 MFMA layout covers these forms only
static bool vec_mfma_form(const skfa_plan* P, bool fast, int nqg, int nqm) {
    return fast && P->C == 3 && P->n == 10 && nqg == 4 && nqm >= 1 && nqm <= 4;
}

static bool mfma_gather(const skfa_plan* P, int nqg, int nqm) {
    return P->tune_mfma > 0 && nqg <= 4 && nqm <= 4;
}

// element tables are automatic for 4-node elements at one row per lane;
// with the rows split over lanes (small meshes) the quadrature gather's
// incidence split avoids the table gather's repeated element rows
static bool table_gather(const skfa_plan* P) {
    return P->tune_table < 0 ? (P->n == 4 && P->min_split == 1) : P->tune_table != 0;
}

// Payload layout of a symmetric scalar form: the MFMA and element-table
// gathers read PL_SYM; the quadrature gather reads the compact layouts
static int sym_layout(const skfa_plan* P, int nqg, int nqm, bool f32, bool mass) {
    if (mfma_gather(P, nqg, nqm) || table_gather(P)) return PL_SYM;
    if (f32) return PL_SYM32;
    return mass ? PL_SYM : PL_SYM6;
}

// The load vector rides in the gather when the element-table gather of
// 4-node scalar fields runs (the choice launch_gather makes).
static bool fused_load_path(const skfa_plan* P, const int32_t* gconn, int32_t per_qp, const double* c_tensor,
                            int32_t flags, int32_t nqg, int32_t nqm, const double* tables, int32_t nq) {
    const bool fast = gconn == nullptr && per_qp == 0;
    const bool diag3 = (flags & SKFA_FLAG_VECTOR3_DIAGONAL) != 0;
    const bool sym = P->C == 1 && (c_tensor == nullptr || (flags & SKFA_FLAG_SYMMETRIC_TENSOR));
    return fast && !diag3 && sym && P->n == 4 && table_gather(P) && !mfma_gather(P, nqg, nqm) && tables != nullptr &&
           nq >= 1 && nq <= 64;
}

// out[map[k]] = src[k] for map[k] >= 0
__global__ void k_scatter_mapped(const double* __restrict__ src, const int32_t* __restrict__ map, int64_t n,
                                 double* __restrict__ out) {
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x) {
        const int32_t j = map[k];
        if (j >= 0) out[j] = src[k];
    }
}

// Scratch of the element-constant path: the per-element payload (E x
// stride doubles) and the element tables of the table gather (FP64, FP32).
struct Workspace {
    double* payload = nullptr;
    double* ntab = nullptr;
    float* ntab32 = nullptr;
};

template <int NL, int C, bool F32>
static int launch_gather(skfa_plan* P, const GatherArgs& A, bool sym, const Workspace& W, size_t lds,
                         hipStream_t s) {
    if constexpr (C == 3) {
        if constexpr (NL == 10) {
            if (P->row_lanes == 4) return launch_vec10_mfma<F32>(P, A, lds, s);
            if (A.Qg == 4) return launch_vec10<F32>(P, A, lds, s);
        }
        return launch_gather_pl<NL, 3, PL_VEC, F32>(P, A, lds, s);
    }
    // MFMA row contraction: 1 = every wave; 2 = waves of at most 16 rows (one
    // MFMA row group, the long-row waves where column-split lanes would
    // recompute whole element rows), the rest by the column-split gather
    const int mode = P->tune_mfma < 0 ? 0 : P->tune_mfma;
    if (!A.plan.diag3 && mfma_gather(P, A.Qg, A.Qm)) {
        GatherArgs M = A;
        if (mode == 2) M.rmax = 16;
        if (int rc = sym ? launch_mfma<NL, PL_SYM, F32>(P, M, s) : launch_mfma<NL, PL_FULL, F32>(P, M, s)) return rc;
        if (mode == 1) return kOk;
        GatherArgs Rest = A;
        Rest.rmin = 17;
        return sym ? launch_gather_pl<NL, 1, PL_SYM, F32>(P, Rest, lds, s)
                   : launch_gather_pl<NL, 1, PL_FULL, F32>(P, Rest, lds, s);
    }
    // element tables win for 4-node elements (few table rows, long column
    // split); the quadrature gather with uniform table reads wins for P2
    if (sym && table_gather(P)) {
        // element tables from the reduced quadrature tables, then the
        // table-driven gather
        // (the payload pass built the element tables from the reduced
        // quadrature tables into W.ntab / W.ntab32)
        SKFA_REQUIRE(W.ntab != nullptr && W.ntab32 != nullptr, "element-table buffers are missing");
        GatherArgs T = A;
        T.tab = W.ntab;
        T.tab32 = W.ntab32;
        if constexpr (NL == 4) {
            if (A.load_out != nullptr)
                return A.has_mass ? launch_table<NL, F32, true, true>(P, T, s)
                                  : launch_table<NL, F32, false, true>(P, T, s);
        }
        return A.has_mass ? launch_table<NL, F32, true>(P, T, s) : launch_table<NL, F32, false>(P, T, s);
    }
    if (sym) {
        const int lay = sym_layout(P, A.Qg, A.Qm, F32, A.has_mass != 0);
        if constexpr (F32) {
            if (lay == PL_SYM32) return launch_gather_pl<NL, 1, PL_SYM32, true>(P, A, lds, s);
        } else {
            if (lay == PL_SYM6) return launch_gather_pl<NL, 1, PL_SYM6, false>(P, A, lds, s);
        }
        return launch_gather_pl<NL, 1, PL_SYM, F32>(P, A, lds, s);
    }
    return launch_gather_pl<NL, 1, PL_FULL, F32>(P, A, lds, s);
}

template <int NL, int C, bool ISO>
static int launch_general(const skfa_plan* P, const H1Args& A, size_t lds, hipStream_t s) {
    const dim3 grid = wave_grid(P), block(64);
    return with_record_layout<NL>(P, [&](auto rb, auto ow) {
        constexpr int RB = decltype(rb)::value, OFFW = decltype(ow)::value;
        k_assemble_h1_general<NL, C, ISO, RB, OFFW><<<grid, block, lds, s>>>(A);
        SKFA_LAUNCH_CHECK();
        return kOk;
    });
}

template <int NL, int C>
static int dispatch_nl(skfa_plan* P, bool fast, bool iso, bool f32, bool sym, const GatherArgs& G,
                       const H1Args& H, const Workspace& W, size_t lds_fast, size_t lds_general, hipStream_t s) {
    if (fast)
        return f32 ? launch_gather<NL, C, true>(P, G, sym, W, lds_fast, s)
                   : launch_gather<NL, C, false>(P, G, sym, W, lds_fast, s);
    return iso ? launch_general<NL, C, true>(P, H, lds_general, s) : launch_general<NL, C, false>(P, H, lds_general, s);
}

// The bilinear entry points.  With `scratch` null the plan owns the
// payload and element-table buffers (allocated on first use); otherwise
// they are carved from `scratch` (scratch_bytes long) and nothing is
// allocated.
// ------------------------------------------------------------------ no-allocation scratch
// Device memory of skfa_assemble_bilinear_noalloc: a file-scope array that
// every GPU the library runs on holds until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns, looked up once per device with that device
// current.  Calls that overlap on one GPU share the array and race.
constexpr size_t kScratchBytes = size_t(96) << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mutex;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

static int device_scratch(unsigned char** out) {
    int dev = 0;
    SKFA_HIP(hipGetDevice(&dev));
    SKFA_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        SKFA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        SKFA_REQUIRE(ptr != nullptr, "scratch symbol has no device address");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return kOk;
}

static int assemble_bilinear_ws(skfa_plan* P, const double* xyz, const int32_t* vconn, const int32_t* gconn,
                                int32_t gnodes, const double* tables, int32_t nq, const double* rtab,
                                const float* rtab32, int32_t nqg, int32_t nqm, const double* c_value,
                                const double* c_gradient, const double* c_symgrad, const double* c_divergence,
                                const double* c_tensor, int32_t per_qp, int32_t precision, int32_t flags,
                                double* values, void* stream, unsigned char* scratch, size_t scratch_bytes,
                                const double* load_f = nullptr, double* load_out = nullptr,
                                bool* load_fused = nullptr, const int32_t* value_map = nullptr,
                                const int32_t* row_map = nullptr) {
    if (load_fused) *load_fused = false;
    SKFA_REQUIRE(P != nullptr, "plan is null");
    SKFA_REQUIRE(P->n == 4 || P->n == 10, "bilinear assembly supports 4- and 10-node tetrahedra");
    SKFA_REQUIRE(P->C == 1 || P->C == 3, "bilinear assembly supports 1 or 3 components on tetrahedra");
    SKFA_REQUIRE(values != nullptr || P->nnz() == 0, "values pointer is null");
    SKFA_REQUIRE(xyz != nullptr && vconn != nullptr, "geometry pointer is null");
    SKFA_REQUIRE(c_value || c_gradient || c_symgrad || c_divergence || c_tensor,
                 "at least one coefficient is required");
    SKFA_REQUIRE(!(c_tensor && P->C != 1), "anisotropic gradient tensor currently requires a scalar field");
    const bool iso = gconn != nullptr;
    const bool fast = !iso && per_qp == 0;
    SKFA_REQUIRE(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (FP32 products)");
    SKFA_REQUIRE(!(precision == 1 && !fast),
                 "FP32 products are available for affine elements with element-constant coefficients");
    SKFA_REQUIRE(!fast || (rtab != nullptr && nqg >= 1 && nqm >= 1 && nqg <= 64 && nqm <= 64),
                 "reduced tables are required for element-constant coefficients");
    SKFA_REQUIRE(!(precision == 1 && rtab32 == nullptr), "FP32 tables are null");
    SKFA_REQUIRE(fast || (tables != nullptr && nq >= 1 && nq <= 4096), "tables / quadrature size invalid");
    SKFA_REQUIRE(!iso || (gnodes >= 4 && gnodes <= 64), "geometry node count out of range");
    const bool diag3 = (flags & SKFA_FLAG_VECTOR3_DIAGONAL) != 0;
    SKFA_REQUIRE(!diag3 || (P->C == 1 && fast && !c_symgrad && !c_divergence && !c_tensor),
                 "diagonal 3-component output needs a scalar plan, affine elements, element-constant value / "
                 "gradient coefficients only");
    SKFA_REQUIRE(P->row_lanes != 4 || vec_mfma_form(P, fast, nqg, nqm),
                 "a plan with the 4-lane vector MFMA layout assembles affine elements with element-constant "
                 "coefficients, a 4-point gradient rule and at most 4 mass points");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->nnz() == 0 || P->E == 0) {
        if (P->nnz()) SKFA_HIP(hipMemsetAsync(values, 0, sizeof(double) * P->nnz() * (diag3 ? 9 : 1), s));
        return kOk;
    }
    GatherArgs G{};
    H1Args H{};
    Workspace W{};
    bool sym = false;
    size_t lds_fast = 0, lds_general = size_t(P->lds_bytes);
    if (fast) {
        // D is symmetric without a tensor term, or with a tensor the caller
        // declares symmetric
        sym = P->C == 1 && (c_tensor == nullptr || (flags & SKFA_FLAG_SYMMETRIC_TENSOR));
        // the FP32-products vector-P2 gather reads a float payload (16 floats)
        const bool vec32 = P->C == 3 && P->n == 10 && nqg == 4 && precision == 1;
        const int stride = P->C == 1 ? (sym ? 8 : 10) : (vec32 ? 8 : 14);
        // the table gather runs for symmetric scalar forms unless the MFMA
        // gather takes them (launch_gather)
        const bool table = P->C == 1 && sym && table_gather(P) && !(!diag3 && mfma_gather(P, nqg, nqm));
        const size_t nt = size_t(P->n) * P->n * kNK;
        if (scratch == nullptr) {
            if (int rc = ensure_payload(P, stride)) return rc;
            if (table && P->ntab.n < nt)
                if (int rc = P->ntab.alloc(nt)) return rc;
            if (table && P->ntab32.n < nt)
                if (int rc = P->ntab32.alloc(nt)) return rc;
            W.payload = P->payload.p;
            W.ntab = P->ntab.p;
            W.ntab32 = P->ntab32.p;
        } else {
            // 256-byte aligned pieces: payload | FP64 tables | FP32 tables
            auto up = [](size_t b) { return (b + 255) / 256 * 256; };
            const size_t pay = up(size_t(P->E) * stride * sizeof(double));
            const size_t need = pay + (table ? up(nt * sizeof(double)) + up(nt * sizeof(float)) : 0);
            if (need > scratch_bytes) {
                set_error("the element payload of " + std::to_string(P->E) + " elements needs " +
                          std::to_string(need) + " bytes, more than the " + std::to_string(scratch_bytes) +
                          "-byte scratch of the no-allocation entry point");
                return kCapacity;
            }
            W.payload = reinterpret_cast<double*>(scratch);
            if (table) {
                W.ntab = reinterpret_cast<double*>(scratch + pay);
                W.ntab32 = reinterpret_cast<float*>(scratch + pay + up(nt * sizeof(double)));
            }
        }
        PayloadArgs pa{};
        pa.xyz = reinterpret_cast<const double4*>(xyz);
        pa.vconn = vconn;
        pa.E = P->E;
        pa.cm = c_value; pa.cg = c_gradient; pa.cs = c_symgrad; pa.cl = c_divergence; pa.cA = c_tensor;
        pa.out = W.payload;
        if (table) {
            pa.rtab = rtab;
            pa.tab_n = P->n;
            pa.tab_qg = nqg;
            pa.tab_qm = nqm;
            pa.ntab = W.ntab;
            pa.ntab32 = W.ntab32;
        }
        const int lay = sym ? sym_layout(P, nqg, nqm, precision == 1, c_value != nullptr) : PL_FULL;
        if (P->C == 1 && lay == PL_SYM32)
            k_payload_h1<1, true, SYMOUT_F32><<<elem_grid(P->E), 256, 0, s>>>(pa);
        else if (P->C == 1 && lay == PL_SYM6)
            k_payload_h1<1, true, SYMOUT_F64_NOMASS><<<elem_grid(P->E), 256, 0, s>>>(pa);
        else if (P->C == 1 && sym)
            k_payload_h1<1, true><<<elem_grid(P->E), 256, 0, s>>>(pa);
        else if (P->C == 1)
            k_payload_h1<1, false><<<elem_grid(P->E), 256, 0, s>>>(pa);
        else if (vec32)
            k_payload_h1<3, false, SYMOUT_F32><<<elem_grid(P->E), 256, 0, s>>>(pa);
        else
            k_payload_h1<3, false><<<elem_grid(P->E), 256, 0, s>>>(pa);
        SKFA_LAUNCH_CHECK();
        G.plan = plan_args(P);
        G.plan.diag3 = diag3 ? 1 : 0;
        G.payload = W.payload;
        G.tab = rtab;
        G.tab32 = rtab32;
        G.Qg = nqg;
        G.Qm = nqm;
        if ((flags & SKFA_FLAG_BARYCENTRIC_TABLE) && P->n == 10) {
            // the 16 x Qg barycentric coefficients follow wg | dg | wm | pm
            const size_t off = size_t(nqg) + size_t(nqg) * 30 + nqm + size_t(nqm) * 10;
            G.ctab = rtab + off;
            G.ctab32 = rtab32 ? rtab32 + off : nullptr;
            // then the 10 x 10 x 8 pair table, at a multiple of 4 entries
            const size_t poff = (off + size_t(16) * nqg + 3) & ~size_t(3);
            G.qtab = rtab + poff;
            G.qtab32 = rtab32 ? rtab32 + poff : nullptr;
        }
        G.has_mass = c_value != nullptr;
        G.values = values;
        // the constant-source load vector rides along with the element-table
        // gather of 4-node scalar fields (|det J| in the payload's last slot)
        if (load_out != nullptr &&
            fused_load_path(P, gconn, per_qp, c_tensor, flags, nqg, nqm, tables, nq)) {
            G.load_tables = tables;
            G.load_q = nq;
            G.load_f = load_f;
            G.load_out = load_out;
            G.value_map = value_map;
            G.row_map = row_map;
            if (load_fused) *load_fused = true;
        }
        const size_t tab_len = size_t(nqg) + size_t(nqg) * P->n * 3 + nqm + size_t(nqm) * P->n;
        lds_fast = size_t(P->lds_bytes) + tab_len * (precision == 1 ? 4 : 8);
    } else {
        H.plan = plan_args(P);
        H.xyz = reinterpret_cast<const double4*>(xyz);
        H.vconn = vconn;
        H.gconn = gconn;
        H.gn = iso ? gnodes : 4;
        H.tables = tables;
        H.Q = nq;
        H.cm = c_value; H.cg = c_gradient; H.cs = c_symgrad; H.cl = c_divergence; H.cA = c_tensor;
        H.per_qp = per_qp;
        H.values = values;
    }
    SKFA_REQUIRE(lds_fast <= 160 * 1024 && lds_general <= 160 * 1024, "LDS request exceeds 160 KiB");
    const bool f32 = precision == 1;
    if (P->n == 4)
        return P->C == 1 ? dispatch_nl<4, 1>(P, fast, iso, f32, sym, G, H, W, lds_fast, lds_general, s)
                         : dispatch_nl<4, 3>(P, fast, iso, f32, sym, G, H, W, lds_fast, lds_general, s);
    return P->C == 1 ? dispatch_nl<10, 1>(P, fast, iso, f32, sym, G, H, W, lds_fast, lds_general, s)
                     : dispatch_nl<10, 3>(P, fast, iso, f32, sym, G, H, W, lds_fast, lds_general, s);
}

}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_assemble_bilinear(skfa_plan* P, const double* xyz, int64_t npoints, const int32_t* vconn,
                           const int32_t* gconn, int32_t gnodes, const double* tables, int32_t nq,
                           const double* rtab, const float* rtab32, int32_t nqg, int32_t nqm,
                           const double* c_value, const double* c_gradient, const double* c_symgrad,
                           const double* c_divergence, const double* c_tensor, int32_t per_qp, int32_t precision,
                           int32_t flags, double* values, void* stream) {
    (void)npoints;
    return assemble_bilinear_ws(P, xyz, vconn, gconn, gnodes, tables, nq, rtab, rtab32, nqg, nqm, c_value,
                                c_gradient, c_symgrad, c_divergence, c_tensor, per_qp, precision, flags, values,
                                stream, nullptr, 0);
}

int skfa_assemble_bilinear_load(skfa_plan* P, const double* xyz, int64_t npoints, const int32_t* vconn,
                                const int32_t* gconn, int32_t gnodes, const double* tables, int32_t nq,
                                const double* rtab, const float* rtab32, int32_t nqg, int32_t nqm,
                                const double* c_value, const double* c_gradient, const double* c_symgrad,
                                const double* c_divergence, const double* c_tensor, int32_t per_qp,
                                int32_t precision, int32_t flags, double* values, const double* f, double* load,
                                const int32_t* value_map, const int32_t* row_map, void* stream) {
    (void)npoints;
    SKFA_REQUIRE(P != nullptr, "plan is null");
    SKFA_REQUIRE(f != nullptr && load != nullptr, "load source or output is null");
    SKFA_REQUIRE(P->C == 1 && !(flags & SKFA_FLAG_VECTOR3_DIAGONAL), "the fused load vector is for scalar fields");
    SKFA_REQUIRE(tables != nullptr && nq >= 1, "basis tables are required for the load vector");
    SKFA_REQUIRE((value_map == nullptr) == (row_map == nullptr), "give both output maps or neither");
    const bool mapped = value_map != nullptr;
    const bool fusable = fused_load_path(P, gconn, per_qp, c_tensor, flags, nqg, nqm, tables, nq);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (mapped && (!fusable || P->nnz() == 0 || P->E == 0)) {
        // full-size results in plan-owned buffers, then through the maps
        if (P->full_values.n < size_t(std::max<int64_t>(P->nnz(), 1)))
            if (int rc = P->full_values.alloc(size_t(std::max<int64_t>(P->nnz(), 1)))) return rc;
        if (P->full_load.n < size_t(std::max<int64_t>(P->nnodes, 1)))
            if (int rc = P->full_load.alloc(size_t(std::max<int64_t>(P->nnodes, 1)))) return rc;
        if (int rc = assemble_bilinear_ws(P, xyz, vconn, gconn, gnodes, tables, nq, rtab, rtab32, nqg, nqm, c_value,
                                          c_gradient, c_symgrad, c_divergence, c_tensor, per_qp, precision, flags,
                                          P->full_values.p, stream, nullptr, 0))
            return rc;
        if (int rc = skfa_assemble_linear(P, xyz, vconn, gconn, gnodes, tables, nq, f, 0, nullptr, 0,
                                          P->full_load.p, stream))
            return rc;
        if (P->nnz())
            k_scatter_mapped<<<unsigned(std::min<int64_t>((P->nnz() + 255) / 256, 65536)), 256, 0, s>>>(
                P->full_values.p, value_map, P->nnz(), values);
        if (P->nnodes)
            k_scatter_mapped<<<unsigned(std::min<int64_t>((P->nnodes + 255) / 256, 65536)), 256, 0, s>>>(
                P->full_load.p, row_map, P->nnodes, load);
        SKFA_LAUNCH_CHECK();
        return kOk;
    }
    bool fused = false;
    if (int rc = assemble_bilinear_ws(P, xyz, vconn, gconn, gnodes, tables, nq, rtab, rtab32, nqg, nqm, c_value,
                                      c_gradient, c_symgrad, c_divergence, c_tensor, per_qp, precision, flags,
                                      values, stream, nullptr, 0, f, load, &fused, value_map, row_map))
        return rc;
    if (fused) return kOk;
    // every other gather: the load vector by the linear-form kernels
    return skfa_assemble_linear(P, xyz, vconn, gconn, gnodes, tables, nq, f, 0, nullptr, 0, load, stream);
}

int skfa_assemble_bilinear_noalloc(skfa_plan* P, const double* xyz, int64_t npoints, const int32_t* vconn,
                                   const int32_t* gconn, int32_t gnodes, const double* tables, int32_t nq,
                                   const double* rtab, const float* rtab32, int32_t nqg, int32_t nqm,
                                   const double* c_value, const double* c_gradient, const double* c_symgrad,
                                   const double* c_divergence, const double* c_tensor, int32_t per_qp,
                                   int32_t precision, int32_t flags, double* values, void* stream) {
    (void)npoints;
    unsigned char* scratch_ptr = nullptr;
    if (int rc = device_scratch(&scratch_ptr)) return rc;
    return assemble_bilinear_ws(P, xyz, vconn, gconn, gnodes, tables, nq, rtab, rtab32, nqg, nqm, c_value,
                                c_gradient, c_symgrad, c_divergence, c_tensor, per_qp, precision, flags, values,
                                stream, scratch_ptr, kScratchBytes);
}

int skfa_assemble_tet_n1(skfa_plan* P, const double* xyz, const int32_t* vconn, const double* tables,
                         int32_t nq, const double* kappa, const double* mu, int32_t per_qp, int32_t include_mass,
                         int32_t include_curl, double* values, void* stream) {
    SKFA_REQUIRE(P != nullptr && P->n == 6 && P->C == 1, "plan must be an edge plan (6 local edges, 1 component)");
    SKFA_REQUIRE(xyz && vconn && tables, "geometry pointer is null");
    SKFA_REQUIRE(!include_mass || kappa, "mass coefficient is null");
    SKFA_REQUIRE(!include_curl || mu, "curl coefficient is null");
    SKFA_REQUIRE(nq >= 1 && nq <= 64, "quadrature size out of range");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->nnz() == 0 || P->E == 0) {
        if (P->nnz()) SKFA_HIP(hipMemsetAsync(values, 0, sizeof(double) * P->nnz(), s));
        return kOk;
    }
    // element-constant coefficients with the 4-point rule: the corner form
    // (no payload pass); otherwise payload pass + gather
    const bool corners = !per_qp && nq == 4 && P->rec_bytes == 16 && (P->off_bytes == 1 || P->off_bytes == 2);
    GatherN1Args A{};
    A.plan = plan_args(P);
    A.xyz = reinterpret_cast<const double4*>(xyz);
    A.vconn = vconn;
    A.kappa = kappa;
    A.mu = mu;
    A.tab = tables;
    A.Q = nq;
    A.include_mass = include_mass;
    A.include_curl = include_curl;
    A.values = values;
    const size_t lds = size_t(P->lds_bytes) + size_t(nq + nq * 18 + 18) * 8;
    const dim3 grid = wave_grid(P), block(64);
    if (corners) {
        if (P->off_bytes == 1)
            k_gather_n1_corners<16, 1><<<grid, block, lds, s>>>(A);
        else
            k_gather_n1_corners<16, 2><<<grid, block, lds, s>>>(A);
        SKFA_LAUNCH_CHECK();
        return kOk;
    }
    if (int rc = ensure_payload(P, 14)) return rc;
    PayloadArgs pa{};
    pa.xyz = reinterpret_cast<const double4*>(xyz);
    pa.vconn = vconn;
    pa.E = P->E;
    pa.cm = include_mass ? kappa : nullptr;
    pa.cg = include_curl ? mu : nullptr;
    // per-point coefficients: flag the payload pass to leave them out
    pa.cs = per_qp && include_mass ? kappa : nullptr;
    pa.cl = per_qp && include_curl ? mu : nullptr;
    pa.wtab = tables;
    pa.Q = nq;
    pa.out = P->payload.p;
    k_payload_n1<<<elem_grid(P->E), 256, 0, s>>>(pa);
    SKFA_LAUNCH_CHECK();
    A.payload = P->payload.p;
    A.kappa_qp = per_qp && include_mass ? kappa : nullptr;
    A.mu_qp = per_qp && include_curl ? mu : nullptr;
    // 6 edges: 4 + 6 offset bytes (16 B records) with 1- or 2-byte offsets
    if (P->rec_bytes == 16 && P->off_bytes == 1) {
        if (nq == 4)
            k_gather_n1<16, 1, 4><<<grid, block, lds, s>>>(A);
        else
            k_gather_n1<16, 1, 0><<<grid, block, lds, s>>>(A);
    } else if (P->rec_bytes == 16 && P->off_bytes == 2) {
        if (nq == 4)
            k_gather_n1<16, 2, 4><<<grid, block, lds, s>>>(A);
        else
            k_gather_n1<16, 2, 0><<<grid, block, lds, s>>>(A);
    } else {
        set_error("unsupported gather record layout");
        return kUnsupported;
    }
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_assemble_linear(skfa_plan* P, const double* xyz, const int32_t* vconn, const int32_t* gconn,
                         int32_t gnodes, const double* tables, int32_t nq, const double* f, int32_t f_per_qp,
                         const double* G, int32_t G_per_qp, double* out, void* stream) {
    SKFA_REQUIRE(P != nullptr && out != nullptr, "null argument");
    SKFA_REQUIRE(xyz && vconn && tables, "geometry pointer is null");
    SKFA_REQUIRE(f || G, "at least one coefficient is required");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->nrows() == 0) return kOk;
    LinArgs A;
    A.xyz = reinterpret_cast<const double4*>(xyz);
    A.vconn = vconn;
    A.gconn = gconn;
    A.gn = gconn ? gnodes : 4;
    A.tables = tables;
    A.Q = nq;
    A.n = P->n;
    A.C = P->C;
    A.f = f;
    A.f_per_qp = f_per_qp;
    A.G = G;
    A.G_per_qp = G_per_qp;
    A.inc_ptr = P->inc_ptr.p;
    A.inc = P->inc.p;
    A.nnodes = P->nnodes;
    A.out = out;
    const int64_t total = P->nrows();
    unsigned grid = unsigned(std::min<int64_t>((total + 255) / 256, 65536));
    if (!gconn && f && !f_per_qp && !G && P->n <= 16) {
        if (P->elem_det.n < size_t(P->E))
            if (int rc = P->elem_det.alloc(size_t(std::max<int64_t>(P->E, 1)))) return rc;
        k_elem_absdet<<<unsigned(std::min<int64_t>((P->E + 255) / 256, 65536)), 256, 0, s>>>(A.xyz, vconn, P->E,
                                                                                              P->elem_det.p);
        k_linear_const<<<unsigned(std::min<int64_t>((P->nnodes * 16 + 255) / 256, 65536)), 256, 0, s>>>(
            A.inc_ptr, A.inc, P->nnodes, A.C, P->elem_det.p, tables, nq, A.n, A.gn, f, out);
        SKFA_LAUNCH_CHECK();
        return kOk;
    }
    if (gconn)
        k_assemble_linear<true><<<grid, 256, 0, s>>>(A);
    else
        k_assemble_linear<false><<<grid, 256, 0, s>>>(A);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_tabulate_geometry(const double* xyz, const int32_t* vconn, const int32_t* gconn, int32_t gnodes, int64_t E,
                           const double* tables, int32_t nq, int32_t n, double* det, double* tol, double* cond,
                           double* dx, double* x, double* shape, double* grads, void* stream) {
    SKFA_REQUIRE(xyz && tables, "geometry pointer is null");
    SKFA_REQUIRE(gconn || vconn, "connectivity pointer is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (E == 0) return kOk;
    GeoArgs A;
    A.xyz = reinterpret_cast<const double4*>(xyz);
    A.vconn = vconn;
    A.gconn = gconn;
    A.gn = gconn ? gnodes : 4;
    A.tables = tables;
    A.Q = nq;
    A.n = n;
    A.E = E;
    A.det = det; A.tol = tol; A.cond = cond; A.dx = dx; A.x = x; A.shape = shape; A.grads = grads;
    const int64_t total = E * nq;
    unsigned grid = unsigned(std::min<int64_t>((total + 255) / 256, 65536));
    if (gconn)
        k_geometry<true><<<grid, 256, 0, s>>>(A);
    else
        k_geometry<false><<<grid, 256, 0, s>>>(A);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

}  // extern "C"
