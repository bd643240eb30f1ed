// FP64 sparse kernels used by the preconditioned conjugate-gradient solver
// that consumes assembled systems.  The reference has no solver (solver
// policy is external, python/skfemntv/capabilities.py:67); these kernels
// exist for the learned-preconditioner workload built on the assembler.
#include <hip/hip_runtime.h>

#include "skfa_internal.h"
#include "../../include/skfem_amd.h"

namespace skfa {

// y = A x for CSR A; one 64-lane wave per row block of 64/SUB rows, SUB
// lanes per row, deterministic in-row reduction order.
template <int SUB>
__global__ void __launch_bounds__(256) k_spmv(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                              const double* __restrict__ vals, const double* __restrict__ x,
                                              double* __restrict__ y, int64_t nrows) {
    const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t row = gid / SUB;
    const int sub = int(gid % SUB);
    double acc = 0.;
    if (row < nrows) {
        const int64_t b = indptr[row], e = indptr[row + 1];
        for (int64_t k = b + sub; k < e; k += SUB) acc += vals[k] * x[indices[k]];
    }
#pragma unroll
    for (int off = SUB / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, SUB);
    if (row < nrows && sub == 0) y[row] = acc;
}

}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_spmv(const int64_t* indptr, const int32_t* indices, const double* vals, const double* x, double* y,
              int64_t nrows, int32_t lanes_per_row, void* stream) {
    SKFA_REQUIRE(indptr && x && y, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nrows == 0) return kOk;
    const int sub = lanes_per_row;
    const int64_t threads = nrows * sub;
    const unsigned grid = unsigned((threads + 255) / 256);
    switch (sub) {
        case 4: k_spmv<4><<<grid, 256, 0, s>>>(indptr, indices, vals, x, y, nrows); break;
        case 8: k_spmv<8><<<grid, 256, 0, s>>>(indptr, indices, vals, x, y, nrows); break;
        case 16: k_spmv<16><<<grid, 256, 0, s>>>(indptr, indices, vals, x, y, nrows); break;
        default: set_error("lanes_per_row must be 4, 8 or 16"); return kInvalid;
    }
    SKFA_LAUNCH_CHECK();
    return kOk;
}

}  // extern "C"
