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
// bit-reproducible run to run; there are no atomics and no colouring.
#include <hip/hip_runtime.h>

#include <string>

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

struct H1Args {
    const double4* xyz;    // (P) x,y,z,pad
    const int32_t* vconn;  // (E,4) corner vertices
    const int32_t* gconn;  // (E,gn) geometry nodes (isoparametric) or null
    int32_t gn;
    const double* tables;
    const float* tables32;  // FP32 copy of the tables (FP32-products form)
    int32_t Q;
    const double* cm;  // value coefficient
    const double* cg;  // gradient coefficient
    const double* cs;  // symmetric-gradient coefficient
    const double* cl;  // divergence coefficient
    const double* cA;  // gradient tensor (3x3)
    int32_t per_qp;    // coefficients per (e,q) instead of per element
    // plan
    const int64_t* indptr;
    const int64_t* wave_ptr;
    const uint8_t* records;
    const int32_t* bucket_waves;
    int64_t nnodes;
    double* values;
};

__device__ inline void load_coef(const H1Args& A, int64_t idx, Coef& k) {
    k.m = A.cm ? A.cm[idx] : 0.;
    k.g = A.cg ? A.cg[idx] : 0.;
    k.s = A.cs ? A.cs[idx] : 0.;
    k.lam = A.cl ? A.cl[idx] : 0.;
    if (A.cA) {
#pragma unroll
        for (int t = 0; t < 9; ++t) k.A[t] = A.cA[idx * 9 + t];
    } else {
#pragma unroll
        for (int t = 0; t < 9; ++t) k.A[t] = 0.;
    }
}

// D = (1/|det|) * [ dcd (g + s/2) adj adj^T + s/2 adj[:,d] adj[:,c]^T
//                   + lam adj[:,c] adj[:,d]^T + (C==1) adj A adj^T ]
// so that the (a,c),(b,d) entry per quadrature point is
//   w_q * dhat_a^T D dhat_b  (+ mass  w_q |det| m dcd phi_a phi_b).
template <int C>
__device__ inline void build_D(const Geo& G, const Coef& k, int c, int d, bool has_tensor, double* D) {
    const double inv = 1. / fabs(G.det);
    const double* a = G.adj;
    const double iso = (c == d) ? (k.g + 0.5 * k.s) : 0.;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double aat = a[i * 3 + 0] * a[j * 3 + 0] + a[i * 3 + 1] * a[j * 3 + 1] + a[i * 3 + 2] * a[j * 3 + 2];
            D[i * 3 + j] = iso * aat + 0.5 * k.s * a[i * 3 + d] * a[j * 3 + c] + k.lam * a[i * 3 + c] * a[j * 3 + d];
        }
    if (C == 1 && has_tensor) {
        double T[9];  // adj A
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

template <int RB>
struct Record {
    int32_t packed;
    uint8_t off[RB - 4];
};

template <int RB>
__device__ inline void load_record(const uint8_t* __restrict__ base, Record<RB>& r) {
    if constexpr (RB == 8) {
        const uint2 v = *reinterpret_cast<const uint2*>(base);
        r.packed = int32_t(v.x);
        *reinterpret_cast<uint32_t*>(r.off) = v.y;
    } else if constexpr (RB == 16) {
        const uint4 v = *reinterpret_cast<const uint4*>(base);
        r.packed = int32_t(v.x);
        uint32_t* o = reinterpret_cast<uint32_t*>(r.off);
        o[0] = v.y; o[1] = v.z; o[2] = v.w;
    } else {
        const uint4* p = reinterpret_cast<const uint4*>(base);
        uint32_t* o = reinterpret_cast<uint32_t*>(r.off);
        const uint4 v0 = p[0];
        r.packed = int32_t(v0.x);
        o[0] = v0.y; o[1] = v0.z; o[2] = v0.w;
#pragma unroll
        for (int k = 1; k < RB / 16; ++k) {
            const uint4 v = p[k];
            o[4 * k - 1] = v.x; o[4 * k] = v.y; o[4 * k + 1] = v.z; o[4 * k + 2] = v.w;
        }
    }
}

template <int RB, int OFFW>
__device__ inline int rec_offset(const Record<RB>& r, int b) {
    if constexpr (OFFW == 1) return r.off[b];
    else return reinterpret_cast<const uint16_t*>(r.off)[b];
}

// ------------------------------------------------------------------ H1 kernel
// NL: local nodes; C: components; ISO: per-qp Jacobian from gn nodes;
// F32: element products in FP32 (geometry and accumulation stay FP64).
template <int NL, int C, bool ISO, bool F32, int RB, int OFFW>
__global__ void __launch_bounds__(64) k_assemble_h1(H1Args A, int64_t bucket_base) {
    extern __shared__ double span_lds[];
    constexpr int LPR = C * C;
    constexpr int R = 64 / LPR;
    const int w = A.bucket_waves[bucket_base + blockIdx.x];
    const int lane = threadIdx.x;
    const int i = lane / LPR;
    const int cd = lane - i * LPR;
    const int c = cd / C, d = cd - (cd / C) * C;
    const int64_t r0 = int64_t(w) * R;
    const int64_t rn = r0 + i;
    const bool active = (i < R) && (rn < A.nnodes);
    const int64_t rend = (r0 + R < A.nnodes) ? r0 + R : A.nnodes;
    const int64_t lo = A.indptr[r0], hi = A.indptr[rend];
    const int64_t span = (hi - lo) * LPR;
    for (int64_t k = lane; k < span; k += 64) span_lds[k] = 0.;
    int64_t rb = 0;
    if (active) {
        const int64_t len = A.indptr[rn + 1] - A.indptr[rn];
        rb = int64_t(LPR) * (A.indptr[rn] - lo) + int64_t(c) * C * len + d;
    }
    __syncthreads();

    const int Q = A.Q;
    const Tables T = tables_at(A.tables, Q, NL, A.gn);
    const float* T32 = A.tables32;
    const bool has_tensor = A.cA != nullptr;
    const bool has_mass = A.cm != nullptr;
    const int64_t s0 = A.wave_ptr[w], s1 = A.wave_ptr[w + 1];
    for (int64_t st = s0; st < s1; ++st) {
        if (!active) continue;
        Record<RB> rec;
        load_record<RB>(A.records + (st * R + i) * RB, rec);
        if (rec.packed < 0) continue;
        const int64_t e = rec.packed >> 4;
        const int a = rec.packed & 15;
        const int32_t* c4 = A.vconn + e * 4;
        Geo G;
        double Jm[9];
        Coef k;
        double D[9];
        double mfac = 0.;
        if (!ISO) {
            affine_jacobian(A.xyz, c4, Jm);
            adjugate(Jm, G);
            if (!A.per_qp) {
                load_coef(A, e, k);
                build_D<C>(G, k, c, d, has_tensor, D);
                mfac = (c == d) ? fabs(G.det) * k.m : 0.;
            }
        }
        if constexpr (F32) {
            float K[NL];
#pragma unroll
            for (int b = 0; b < NL; ++b) K[b] = 0.f;
            const float* w32 = T32;
            const float* phi32 = w32 + Q;
            const float* dphi32 = phi32 + Q * NL;
            for (int q = 0; q < Q; ++q) {
                if (ISO || A.per_qp) {
                    if (ISO) {
                        iso_jacobian(A.xyz, A.gconn + e * A.gn, A.gn, T.gdphi + q * A.gn * 3, Jm);
                        adjugate(Jm, G);
                    }
                    load_coef(A, A.per_qp ? e * Q + q : e, k);
                    build_D<C>(G, k, c, d, has_tensor, D);
                    mfac = (c == d) ? fabs(G.det) * k.m : 0.;
                }
                float Df[9];
#pragma unroll
                for (int t = 0; t < 9; ++t) Df[t] = float(D[t]);
                const float wq = w32[q];
                const float ga0 = dphi32[(q * NL + a) * 3 + 0], ga1 = dphi32[(q * NL + a) * 3 + 1],
                            ga2 = dphi32[(q * NL + a) * 3 + 2];
                const float u0 = wq * (ga0 * Df[0] + ga1 * Df[3] + ga2 * Df[6]);
                const float u1 = wq * (ga0 * Df[1] + ga1 * Df[4] + ga2 * Df[7]);
                const float u2 = wq * (ga0 * Df[2] + ga1 * Df[5] + ga2 * Df[8]);
#pragma unroll
                for (int b = 0; b < NL; ++b) {
                    const float* gb = dphi32 + (q * NL + b) * 3;
                    K[b] += u0 * gb[0] + u1 * gb[1] + u2 * gb[2];
                }
                if (has_mass) {
                    const float v = wq * float(mfac) * phi32[q * NL + a];
#pragma unroll
                    for (int b = 0; b < NL; ++b) K[b] += v * phi32[q * NL + b];
                }
            }
#pragma unroll
            for (int b = 0; b < NL; ++b) span_lds[rb + int64_t(C) * rec_offset<RB, OFFW>(rec, b)] += double(K[b]);
        } else {
            double K[NL];
#pragma unroll
            for (int b = 0; b < NL; ++b) K[b] = 0.;
            for (int q = 0; q < Q; ++q) {
                if (ISO || A.per_qp) {
                    if (ISO) {
                        iso_jacobian(A.xyz, A.gconn + e * A.gn, A.gn, T.gdphi + q * A.gn * 3, Jm);
                        adjugate(Jm, G);
                    }
                    load_coef(A, A.per_qp ? e * Q + q : e, k);
                    build_D<C>(G, k, c, d, has_tensor, D);
                    mfac = (c == d) ? fabs(G.det) * k.m : 0.;
                }
                const double wq = T.w[q];
                const double* ga = T.dphi + (q * NL + a) * 3;
                const double u0 = wq * (ga[0] * D[0] + ga[1] * D[3] + ga[2] * D[6]);
                const double u1 = wq * (ga[0] * D[1] + ga[1] * D[4] + ga[2] * D[7]);
                const double u2 = wq * (ga[0] * D[2] + ga[1] * D[5] + ga[2] * D[8]);
#pragma unroll
                for (int b = 0; b < NL; ++b) {
                    const double* gb = T.dphi + (q * NL + b) * 3;
                    K[b] += u0 * gb[0] + u1 * gb[1] + u2 * gb[2];
                }
                if (has_mass) {
                    const double v = wq * mfac * T.phi[q * NL + a];
#pragma unroll
                    for (int b = 0; b < NL; ++b) K[b] += v * T.phi[q * NL + b];
                }
            }
#pragma unroll
            for (int b = 0; b < NL; ++b) span_lds[rb + int64_t(C) * rec_offset<RB, OFFW>(rec, b)] += K[b];
        }
    }
    __syncthreads();
    double* out = A.values + lo * LPR;
    for (int64_t k2 = lane; k2 < span; k2 += 64) out[k2] = span_lds[k2];
}

// ------------------------------------------------------------------ N1 kernel
// Rows are global edges; incidence (e, local edge a); sign of local edge l
// is +1 iff its first vertex has the smaller global id
// (edge_topology.cpp:71).  Per element, with adj/det from the corners:
//   mass  K_ab += s_a s_b sum_q w_q kappa phi_aq^T (adj adj^T/|det|) phi_bq
//   curl  K_ab += s_a s_b (sum_q w_q mu_q) curl_a^T (J^T J/|det|) curl_b
// which is the covariant Piola mapping of hcurl_basis.py:355-364 integrated
// as in integrate_tet_n1 (edge_topology.cpp:196-231).
struct N1Args {
    const double4* xyz;
    const int32_t* vconn;
    const double* tables;  // w[Q] | phi[Q][6][3] | curl[6][3]
    int32_t Q;
    const double* kappa;   // mass coefficient
    const double* mu;      // curl coefficient
    int32_t per_qp;
    int32_t include_mass, include_curl;
    const int64_t* indptr;
    const int64_t* wave_ptr;
    const uint8_t* records;
    const int32_t* bucket_waves;
    int64_t nnodes;
    double* values;
};

__constant__ int kN1Edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

template <int RB, int OFFW>
__global__ void __launch_bounds__(64) k_assemble_n1(N1Args A, int64_t bucket_base) {
    extern __shared__ double span_lds[];
    const int w = A.bucket_waves[bucket_base + blockIdx.x];
    const int lane = threadIdx.x;
    const int64_t r0 = int64_t(w) * 64;
    const int64_t rn = r0 + lane;
    const bool active = rn < A.nnodes;
    const int64_t rend = (r0 + 64 < A.nnodes) ? r0 + 64 : A.nnodes;
    const int64_t lo = A.indptr[r0], hi = A.indptr[rend];
    const int64_t span = hi - lo;
    for (int64_t k = lane; k < span; k += 64) span_lds[k] = 0.;
    const int64_t rb = active ? A.indptr[rn] - lo : 0;
    __syncthreads();
    const int Q = A.Q;
    const double* W = A.tables;
    const double* phi = W + Q;           // [Q][6][3]
    const double* curl = phi + Q * 18;   // [6][3]
    double wsum = 0.;
    for (int q = 0; q < Q; ++q) wsum += W[q];
    const int64_t s0 = A.wave_ptr[w], s1 = A.wave_ptr[w + 1];
    for (int64_t st = s0; st < s1; ++st) {
        if (!active) continue;
        Record<RB> rec;
        load_record<RB>(A.records + (st * 64 + lane) * RB, rec);
        if (rec.packed < 0) continue;
        const int64_t e = rec.packed >> 4;
        const int a = rec.packed & 15;
        const int32_t* c4 = A.vconn + e * 4;
        int32_t vid[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) vid[t] = c4[t];
        double Jm[9];
        Geo G;
        affine_jacobian(A.xyz, c4, Jm);
        adjugate(Jm, G);
        const double inv = 1. / fabs(G.det);
        double Dm[9], Dc[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                Dm[i * 3 + j] = inv * (G.adj[i * 3 + 0] * G.adj[j * 3 + 0] + G.adj[i * 3 + 1] * G.adj[j * 3 + 1] +
                                       G.adj[i * 3 + 2] * G.adj[j * 3 + 2]);
                Dc[i * 3 + j] = inv * (Jm[0 * 3 + i] * Jm[0 * 3 + j] + Jm[1 * 3 + i] * Jm[1 * 3 + j] +
                                       Jm[2 * 3 + i] * Jm[2 * 3 + j]);
            }
        double sgn[6];
#pragma unroll
        for (int l = 0; l < 6; ++l) sgn[l] = vid[kN1Edge[l][0]] < vid[kN1Edge[l][1]] ? 1. : -1.;
        double K[6];
#pragma unroll
        for (int b = 0; b < 6; ++b) K[b] = 0.;
        double cw = 0.;
        if (A.include_mass) {
            for (int q = 0; q < Q; ++q) {
                const double kq = A.per_qp ? A.kappa[e * Q + q] : A.kappa[e];
                const double* pa = phi + (q * 6 + a) * 3;
                const double f = W[q] * kq;
                const double u0 = f * (pa[0] * Dm[0] + pa[1] * Dm[3] + pa[2] * Dm[6]);
                const double u1 = f * (pa[0] * Dm[1] + pa[1] * Dm[4] + pa[2] * Dm[7]);
                const double u2 = f * (pa[0] * Dm[2] + pa[1] * Dm[5] + pa[2] * Dm[8]);
#pragma unroll
                for (int b = 0; b < 6; ++b) {
                    const double* pb = phi + (q * 6 + b) * 3;
                    K[b] += u0 * pb[0] + u1 * pb[1] + u2 * pb[2];
                }
            }
        }
        if (A.include_curl) {
            if (A.per_qp) {
                for (int q = 0; q < Q; ++q) cw += W[q] * A.mu[e * Q + q];
            } else {
                cw = wsum * A.mu[e];
            }
            const double* ca = curl + a * 3;
            const double u0 = cw * (ca[0] * Dc[0] + ca[1] * Dc[3] + ca[2] * Dc[6]);
            const double u1 = cw * (ca[0] * Dc[1] + ca[1] * Dc[4] + ca[2] * Dc[7]);
            const double u2 = cw * (ca[0] * Dc[2] + ca[1] * Dc[5] + ca[2] * Dc[8]);
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                const double* cb = curl + b * 3;
                K[b] += u0 * cb[0] + u1 * cb[1] + u2 * cb[2];
            }
        }
#pragma unroll
        for (int b = 0; b < 6; ++b)
            span_lds[rb + rec_offset<RB, OFFW>(rec, b)] += sgn[a] * sgn[b] * K[b];
    }
    __syncthreads();
    double* out = A.values + lo;
    for (int64_t k2 = lane; k2 < span; k2 += 64) out[k2] = span_lds[k2];
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
template <int NL, int C, bool ISO, bool F32>
static int launch_h1_rb(const skfa_plan* P, const H1Args& A, hipStream_t s) {
    for (int b = 0; b < P->nbuckets; ++b) {
        const int64_t nw = P->bucket_start[b + 1] - P->bucket_start[b];
        if (nw == 0) continue;
        const size_t lds = size_t(P->bucket_lds[b]);
        if (P->rec_bytes == 8 && P->off_bytes == 1)
            k_assemble_h1<NL, C, ISO, F32, 8, 1><<<dim3(unsigned(nw)), dim3(64), lds, s>>>(A, P->bucket_start[b]);
        else if (P->rec_bytes == 16 && P->off_bytes == 1)
            k_assemble_h1<NL, C, ISO, F32, 16, 1><<<dim3(unsigned(nw)), dim3(64), lds, s>>>(A, P->bucket_start[b]);
        else if (P->rec_bytes == 16 && P->off_bytes == 2)
            k_assemble_h1<NL, C, ISO, F32, 16, 2><<<dim3(unsigned(nw)), dim3(64), lds, s>>>(A, P->bucket_start[b]);
        else if (P->rec_bytes == 32 && P->off_bytes == 2)
            k_assemble_h1<NL, C, ISO, F32, 32, 2><<<dim3(unsigned(nw)), dim3(64), lds, s>>>(A, P->bucket_start[b]);
        else {
            set_error("unsupported gather record layout");
            return kUnsupported;
        }
        SKFA_LAUNCH_CHECK();
    }
    return kOk;
}

template <int NL>
static int launch_h1_nl(const skfa_plan* P, const H1Args& A, bool iso, bool f32, hipStream_t s) {
    if (P->C == 1) {
        if (iso) return f32 ? launch_h1_rb<NL, 1, true, true>(P, A, s) : launch_h1_rb<NL, 1, true, false>(P, A, s);
        return f32 ? launch_h1_rb<NL, 1, false, true>(P, A, s) : launch_h1_rb<NL, 1, false, false>(P, A, s);
    }
    if (P->C == 3) {
        if (iso) return f32 ? launch_h1_rb<NL, 3, true, true>(P, A, s) : launch_h1_rb<NL, 3, true, false>(P, A, s);
        return f32 ? launch_h1_rb<NL, 3, false, true>(P, A, s) : launch_h1_rb<NL, 3, false, false>(P, A, s);
    }
    set_error("bilinear assembly supports 1 or 3 components on tetrahedra");
    return kUnsupported;
}

}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_assemble_bilinear(const skfa_plan* P, const double* xyz, int64_t npoints, const int32_t* vconn,
                           const int32_t* gconn, int32_t gnodes, const double* tables, const float* tables32,
                           int32_t nq, const double* c_value, const double* c_gradient, const double* c_symgrad,
                           const double* c_divergence, const double* c_tensor, int32_t per_qp, int32_t precision,
                           double* values, void* stream) {
    SKFA_REQUIRE(P != nullptr, "plan is null");
    SKFA_REQUIRE(values != nullptr || P->nnz() == 0, "values pointer is null");
    SKFA_REQUIRE(xyz != nullptr && vconn != nullptr && tables != nullptr, "geometry pointer is null");
    SKFA_REQUIRE(nq >= 1 && nq <= 4096, "quadrature point count out of range");
    SKFA_REQUIRE(c_value || c_gradient || c_symgrad || c_divergence || c_tensor,
                 "at least one coefficient is required");
    SKFA_REQUIRE(!(c_tensor && P->C != 1), "anisotropic gradient tensor currently requires a scalar field");
    SKFA_REQUIRE(precision == 0 || (precision == 1 && tables32 != nullptr), "invalid precision / FP32 tables");
    const bool iso = gconn != nullptr;
    SKFA_REQUIRE(!iso || (gnodes >= 4 && gnodes <= 64), "geometry node count out of range");
    (void)npoints;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->nnz() == 0 || P->E == 0) {
        if (P->nnz()) SKFA_HIP(hipMemsetAsync(values, 0, sizeof(double) * P->nnz(), s));
        return kOk;
    }
    H1Args A;
    A.xyz = reinterpret_cast<const double4*>(xyz);
    A.vconn = vconn;
    A.gconn = gconn;
    A.gn = iso ? gnodes : 4;
    A.tables = tables;
    A.tables32 = tables32;
    A.Q = nq;
    A.cm = c_value; A.cg = c_gradient; A.cs = c_symgrad; A.cl = c_divergence; A.cA = c_tensor;
    A.per_qp = per_qp;
    A.indptr = P->indptr.p;
    A.wave_ptr = P->wave_ptr.p;
    A.records = P->records.p;
    A.bucket_waves = P->bucket_waves.p;
    A.nnodes = P->nnodes;
    A.values = values;
    const bool f32 = precision == 1;
    switch (P->n) {
        case 4: return launch_h1_nl<4>(P, A, iso, f32, s);
        case 10: return launch_h1_nl<10>(P, A, iso, f32, s);
        default: set_error("bilinear assembly supports 4- and 10-node tetrahedra"); return kUnsupported;
    }
}

int skfa_assemble_tet_n1(const skfa_plan* P, const double* xyz, const int32_t* vconn, const double* tables,
                         int32_t nq, const double* kappa, const double* mu, int32_t per_qp, int32_t include_mass,
                         int32_t include_curl, double* values, void* stream) {
    SKFA_REQUIRE(P != nullptr && P->n == 6 && P->C == 1, "plan must be an edge plan (6 local edges, 1 component)");
    SKFA_REQUIRE(xyz && vconn && tables, "geometry pointer is null");
    SKFA_REQUIRE(!include_mass || kappa, "mass coefficient is null");
    SKFA_REQUIRE(!include_curl || mu, "curl coefficient is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->nnz() == 0 || P->E == 0) {
        if (P->nnz()) SKFA_HIP(hipMemsetAsync(values, 0, sizeof(double) * P->nnz(), s));
        return kOk;
    }
    N1Args A;
    A.xyz = reinterpret_cast<const double4*>(xyz);
    A.vconn = vconn;
    A.tables = tables;
    A.Q = nq;
    A.kappa = kappa;
    A.mu = mu;
    A.per_qp = per_qp;
    A.include_mass = include_mass;
    A.include_curl = include_curl;
    A.indptr = P->indptr.p;
    A.wave_ptr = P->wave_ptr.p;
    A.records = P->records.p;
    A.bucket_waves = P->bucket_waves.p;
    A.nnodes = P->nnodes;
    A.values = values;
    for (int b = 0; b < P->nbuckets; ++b) {
        const int64_t nw = P->bucket_start[b + 1] - P->bucket_start[b];
        if (nw == 0) continue;
        const size_t lds = size_t(P->bucket_lds[b]);
        if (P->rec_bytes == 16 && P->off_bytes == 1)
            k_assemble_n1<16, 1><<<dim3(unsigned(nw)), dim3(64), lds, s>>>(A, P->bucket_start[b]);
        else if (P->rec_bytes == 16 && P->off_bytes == 2)
            k_assemble_n1<16, 2><<<dim3(unsigned(nw)), dim3(64), lds, s>>>(A, P->bucket_start[b]);
        else {
            set_error("unsupported gather record layout");
            return kUnsupported;
        }
        SKFA_LAUNCH_CHECK();
    }
    return kOk;
}

int skfa_assemble_linear(const skfa_plan* P, const double* xyz, const int32_t* vconn, const int32_t* gconn,
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
