// Facet (boundary) linear forms for gfx950 (MI355X).
//
// Replaces LinearFormAssembler::assemble on the facet DOFs of a FacetBasis
// (cpp/src/assembly/linear_form_assembler.cpp:41-137, driven by
// NativeLinearForm over python/skfemntv/basis.py:2442-2685): out[node*C + c]
// = sum over the node's (facet, local node) incidences, in facet order, of
// sum_q w[f,q,c] phi_a(q) with w = dx * f and phi_a the parent element's
// shape function at the facet's quadrature points on its local face.  One
// thread per node row; every row is written, rows off the boundary with 0.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "skfa_internal.h"
#include "skfa_plan.h"
#include "../../include/skfem_amd.h"

namespace skfa {
namespace {

__global__ void __launch_bounds__(256) k_facet_linear(const int64_t* __restrict__ inc_ptr,
                                                     const int32_t* __restrict__ inc, int64_t nnodes, int n,
                                                     int C, int Q, const int32_t* __restrict__ face,
                                                     const double* __restrict__ face_shape,
                                                     const double* __restrict__ w, double* __restrict__ out) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nnodes;
         r += int64_t(gridDim.x) * blockDim.x) {
        double acc[3] = {0., 0., 0.};
        for (int64_t k = inc_ptr[r]; k < inc_ptr[r + 1]; ++k) {
            const int32_t packed = inc[k];
            const int64_t f = packed >> 4;
            const int a = packed & 15;
            const double* sh = face_shape + int64_t(face[f]) * Q * n + a;
            const double* wf = w + f * Q * C;
            for (int q = 0; q < Q; ++q) {
                const double phi = sh[q * n];
                for (int c = 0; c < C; ++c) acc[c] += wf[q * C + c] * phi;
            }
        }
        for (int c = 0; c < C; ++c) out[r * C + c] = acc[c];
    }
}

// Facet mass-type bilinear form: node row r owns the DOF-level span
// [C*C*nptr[r], C*C*nptr[r+1]) (rows r*C + c, columns m*C + d, node-major);
// every (facet, local node a) incidence adds sum_q w[f,q] phi_a phi_b to
// the diagonal component blocks of the row's entries at the facet's nodes
// b, in facet order.  Column positions by binary search in the sorted row.
__global__ void __launch_bounds__(256) k_facet_bilinear(const int64_t* __restrict__ inc_ptr,
                                                       const int32_t* __restrict__ inc,
                                                       const int64_t* __restrict__ nptr,
                                                       const int32_t* __restrict__ nidx,
                                                       const int32_t* __restrict__ conn, int64_t nnodes, int n,
                                                       int C, int Q, const int32_t* __restrict__ face,
                                                       const double* __restrict__ face_shape,
                                                       const double* __restrict__ w, double* __restrict__ values) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nnodes;
         r += int64_t(gridDim.x) * blockDim.x) {
        const int64_t b0 = nptr[r];
        const int64_t len = nptr[r + 1] - b0;
        double* row = values + int64_t(C) * C * b0;
        for (int64_t j = 0; j < len * C * C; ++j) row[j] = 0.;
        for (int64_t k = inc_ptr[r]; k < inc_ptr[r + 1]; ++k) {
            const int32_t packed = inc[k];
            const int64_t f = packed >> 4;
            const int a = packed & 15;
            const double* sh = face_shape + int64_t(face[f]) * Q * n;
            const double* wf = w + f * Q;
            for (int b = 0; b < n; ++b) {
                double v = 0.;
                for (int q = 0; q < Q; ++q) v += wf[q] * sh[q * n + a] * sh[q * n + b];
                const int32_t m = conn[f * n + b];
                int64_t lo = 0, hi = len;  // lower bound of m in the row's columns
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (nidx[b0 + mid] < m)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                for (int c = 0; c < C; ++c) row[(int64_t(c) * len + lo) * C + c] += v;
            }
        }
    }
}

}  // namespace
}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_assemble_facet_linear(const skfa_plan* P, const int32_t* face, const double* face_shape, int32_t nq,
                               const double* weights, double* out, int32_t components, void* stream) {
    SKFA_REQUIRE(P != nullptr && out != nullptr, "null argument");
    SKFA_REQUIRE(P->C == 1, "facet plans are built with one component");
    SKFA_REQUIRE(components >= 1 && components <= 3, "components must be 1, 2 or 3");
    SKFA_REQUIRE(nq >= 1 && nq <= 4096, "quadrature size out of range");
    SKFA_REQUIRE(P->E == 0 || (face && face_shape && weights), "facet arrays are null");
    if (P->nnodes == 0) return kOk;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned grid = unsigned(std::min<int64_t>((P->nnodes + 255) / 256, 65536));
    k_facet_linear<<<grid, 256, 0, s>>>(P->inc_ptr.p, P->inc.p, P->nnodes, P->n, components, nq, face, face_shape,
                                        weights, out);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_assemble_facet_bilinear(const skfa_plan* P, const int32_t* conn, const int32_t* face,
                                 const double* face_shape, int32_t nq, const double* weights, double* values,
                                 void* stream) {
    SKFA_REQUIRE(P != nullptr && (values != nullptr || P->nnz() == 0), "null argument");
    SKFA_REQUIRE(P->C >= 1 && P->C <= 3, "components must be 1, 2 or 3");
    SKFA_REQUIRE(nq >= 1 && nq <= 4096, "quadrature size out of range");
    SKFA_REQUIRE(P->E == 0 || (conn && face && face_shape && weights), "facet arrays are null");
    if (P->nnodes == 0) return kOk;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned grid = unsigned(std::min<int64_t>((P->nnodes + 255) / 256, 65536));
    k_facet_bilinear<<<grid, 256, 0, s>>>(P->inc_ptr.p, P->inc.p, P->indptr.p, P->indices.p, conn, P->nnodes, P->n,
                                          P->C, nq, face, face_shape, weights, values);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

}  // extern "C"
