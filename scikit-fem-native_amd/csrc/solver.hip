// FP64 sparse kernels for solving assembled systems: Dirichlet condensation,
// lower-triangular pattern and transpose, CSR SpMV and a batched
// preconditioned conjugate-gradient driver (Jacobi or factorised sparse
// approximate inverse M^{-1} = G^T G).
//
// The reference has no solver (solver policy is external,
// python/skfemntv/capabilities.py:67); its only solve pattern is the
// condensed direct solve of examples/scikit-fem-gallery/comparison.py:80-87,
// which skfa_csr_condense reproduces on the device.  Several independent
// systems are held as one block-diagonal CSR: system s owns rows
// [sys_ptr[s], sys_ptr[s+1]).  Every reduction has a fixed order, so results
// are bit-reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "skfa_internal.h"
#include "../../include/skfem_amd.h"

namespace skfa {

static unsigned grid_of(int64_t items, int threads, int64_t cap = 65536) {
    int64_t g = (items + threads - 1) / threads;
    return unsigned(std::max<int64_t>(1, std::min<int64_t>(g, cap)));
}

// ------------------------------------------------------------------ SpMV
// y = A x; SUB lanes per row with a fixed in-row reduction order.
template <int SUB, typename V>
__global__ void __launch_bounds__(256) k_spmv(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                              const V* __restrict__ vals, const double* __restrict__ x,
                                              double* __restrict__ y, int64_t nrows) {
    const int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t row = gid / SUB;
    const int sub = int(gid % SUB);
    double acc = 0.;
    if (row < nrows) {
        const int64_t b = indptr[row], e = indptr[row + 1];
        for (int64_t k = b + sub; k < e; k += SUB) acc = fma(double(vals[k]), x[indices[k]], acc);
    }
#pragma unroll
    for (int off = SUB / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, SUB);
    if (row < nrows && sub == 0) y[row] = acc;
}

template <typename V>
static int spmv_launch(const int64_t* indptr, const int32_t* indices, const V* vals, const double* x, double* y,
                       int64_t nrows, int sub, hipStream_t s) {
    if (nrows == 0) return kOk;
    const unsigned grid = unsigned((nrows * sub + 255) / 256);
    switch (sub) {
        case 4: k_spmv<4, V><<<grid, 256, 0, s>>>(indptr, indices, vals, x, y, nrows); break;
        case 8: k_spmv<8, V><<<grid, 256, 0, s>>>(indptr, indices, vals, x, y, nrows); break;
        case 16: k_spmv<16, V><<<grid, 256, 0, s>>>(indptr, indices, vals, x, y, nrows); break;
        default: set_error("lanes_per_row must be 4, 8 or 16"); return kInvalid;
    }
    SKFA_LAUNCH_CHECK();
    return kOk;
}

// ------------------------------------------------------------------ condensation
__global__ void k_keep_count(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                             const uint8_t* __restrict__ keep, int64_t nrows, int32_t* __restrict__ count) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * blockDim.x) {
        int32_t c = 0;
        if (keep[r])
            for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) c += keep[indices[k]] ? 1 : 0;
        count[r] = c;
    }
}

__global__ void k_keep_flag(const uint8_t* __restrict__ keep, int64_t n, int32_t* __restrict__ flag) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += int64_t(gridDim.x) * blockDim.x)
        flag[r] = keep[r] ? 1 : 0;
}

__global__ void k_condense_fill(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                const double* __restrict__ vals, const uint8_t* __restrict__ keep,
                                const int64_t* __restrict__ newid, const int64_t* __restrict__ rowstart, int64_t nrows,
                                int64_t* __restrict__ out_indptr, int32_t* __restrict__ out_indices,
                                double* __restrict__ out_vals) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * blockDim.x) {
        if (!keep[r]) continue;
        const int64_t nr = newid[r];
        int64_t o = rowstart[r];
        out_indptr[nr] = o;
        for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
            const int32_t c = indices[k];
            if (!keep[c]) continue;
            out_indices[o] = int32_t(newid[c]);
            out_vals[o] = vals[k];
            ++o;
        }
    }
}

// ------------------------------------------------------------------ lower pattern / transpose
__global__ void k_lower_count(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices, int64_t nrows,
                              int32_t* __restrict__ count) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * blockDim.x) {
        int32_t c = 0;
        for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) c += indices[k] <= r ? 1 : 0;
        count[r] = c;
    }
}

__global__ void k_lower_fill(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices, int64_t nrows,
                             const int64_t* __restrict__ lptr, int32_t* __restrict__ lidx, int64_t* __restrict__ lsrc) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * blockDim.x) {
        int64_t o = lptr[r];
        for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k)
            if (indices[k] <= r) {
                lidx[o] = indices[k];
                lsrc[o] = k;
                ++o;
            }
    }
}

__global__ void k_col_count(const int32_t* __restrict__ idx, int64_t nnz, int32_t* __restrict__ count) {
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz; k += int64_t(gridDim.x) * blockDim.x)
        atomicAdd(&count[idx[k]], 1);
}

// Transpose: entries of column c become row c, in increasing source-row
// order (deterministic: each output row is filled by a scan over source
// rows in order, one thread per output row segment after a counting pass).
__global__ void k_transpose_fill(const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx, int64_t nrows,
                                 const int64_t* __restrict__ tptr, int32_t* __restrict__ cursor,
                                 int32_t* __restrict__ tidx, int64_t* __restrict__ tsrc) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * blockDim.x) {
        for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k) {
            const int32_t c = idx[k];
            const int32_t slot = atomicAdd(&cursor[c], 1);
            tidx[tptr[c] + slot] = int32_t(r);
            tsrc[tptr[c] + slot] = k;
        }
    }
}

__global__ void k_transpose_sort(const int64_t* __restrict__ tptr, int32_t* __restrict__ tidx,
                                 int64_t* __restrict__ tsrc, int64_t n) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += int64_t(gridDim.x) * blockDim.x) {
        const int64_t b = tptr[r], e = tptr[r + 1];
        for (int64_t i = b + 1; i < e; ++i) {
            const int32_t key = tidx[i];
            const int64_t sv = tsrc[i];
            int64_t j = i - 1;
            while (j >= b && tidx[j] > key) {
                tidx[j + 1] = tidx[j];
                tsrc[j + 1] = tsrc[j];
                --j;
            }
            tidx[j + 1] = key;
            tsrc[j + 1] = sv;
        }
    }
}

template <typename T>
__global__ void k_gather_values(const T* __restrict__ src, const int64_t* __restrict__ map, int64_t n,
                                T* __restrict__ out) {
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x)
        out[k] = src[map[k]];
}

// ------------------------------------------------------------------ PCG
constexpr int kRedThreads = 256;

// Per-system sums sum_i x_i y_i (and optionally sum_i x_i^2): one block per
// system, fixed strided order + tree reduction.
__global__ void __launch_bounds__(kRedThreads) k_sys_dot2(const double* __restrict__ x, const double* __restrict__ y,
                                                          const int64_t* __restrict__ sys_ptr,
                                                          double* __restrict__ out_xy, double* __restrict__ out_xx) {
    __shared__ double s1[kRedThreads], s2[kRedThreads];
    const int s = blockIdx.x;
    double a = 0., b = 0.;
    for (int64_t i = sys_ptr[s] + threadIdx.x; i < sys_ptr[s + 1]; i += kRedThreads) {
        a = fma(x[i], y[i], a);
        if (out_xx) b = fma(x[i], x[i], b);
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s1[threadIdx.x] += s1[threadIdx.x + off];
            s2[threadIdx.x] += s2[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out_xy[s] = s1[0];
        if (out_xx) out_xx[s] = s2[0];
    }
}

__global__ void k_row_system(const int64_t* __restrict__ sys_ptr, int32_t nsys, int32_t* __restrict__ row_sys) {
    const int s = blockIdx.y;
    if (s >= nsys) return;
    for (int64_t i = sys_ptr[s] + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < sys_ptr[s + 1];
         i += int64_t(gridDim.x) * blockDim.x)
        row_sys[i] = s;
}

__global__ void k_diag_inv(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                           const double* __restrict__ vals, int64_t n, double* __restrict__ dinv) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += int64_t(gridDim.x) * blockDim.x) {
        double d = 0.;
        for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k)
            if (indices[k] == r) d = vals[k];
        dinv[r] = 1. / d;
    }
}

__global__ void k_pcg_init(const double* __restrict__ bb, int32_t nsys, int32_t* __restrict__ active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) active[s] = bb[s] > 0. ? 1 : 0;
}

// x += a p, r -= a q for systems still iterating (a = rz / pq).
__global__ void k_pcg_update(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                             const double* __restrict__ q, const double* __restrict__ rz, const double* __restrict__ pq,
                             const int32_t* __restrict__ active, const int32_t* __restrict__ row_sys, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const int s = row_sys[i];
        if (!active[s] || !(pq[s] != 0.)) continue;
        const double a = rz[s] / pq[s];
        x[i] = fma(a, p[i], x[i]);
        r[i] = fma(-a, q[i], r[i]);
    }
}

__global__ void k_jacobi(const double* __restrict__ r, const double* __restrict__ dinv, double* __restrict__ z, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        z[i] = r[i] * dinv[i];
}

// p = z + (rz_new / rz_old) p for active systems
__global__ void k_pcg_direction(double* __restrict__ p, const double* __restrict__ z, const double* __restrict__ rz_new,
                                const double* __restrict__ rz_old, const int32_t* __restrict__ active,
                                const int32_t* __restrict__ row_sys, int64_t n, int first) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const int s = row_sys[i];
        if (!active[s]) continue;
        p[i] = first ? z[i] : fma(rz_new[s] / rz_old[s], p[i], z[i]);
    }
}

// convergence: active[s] = ||r_s||^2 > tol^2 ||b_s||^2; iteration counters
__global__ void k_pcg_check(const double* __restrict__ rr, const double* __restrict__ bb, double tol2, int32_t nsys,
                            int32_t* __restrict__ active, int32_t* __restrict__ iters, int32_t* __restrict__ nactive) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    if (active[s]) {
        iters[s] += 1;
        if (!(rr[s] > tol2 * bb[s])) active[s] = 0;
    }
    if (active[s]) atomicAdd(nactive, 1);
}

}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_spmv(const int64_t* indptr, const int32_t* indices, const double* vals, const double* x, double* y,
              int64_t nrows, int32_t lanes_per_row, void* stream) {
    SKFA_REQUIRE(indptr && x && y, "null argument");
    return spmv_launch(indptr, indices, vals, x, y, nrows, lanes_per_row, static_cast<hipStream_t>(stream));
}

int skfa_csr_condense(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t nrows,
                      const uint8_t* keep, int64_t* newid, int64_t* out_indptr, int32_t* out_indices,
                      double* out_vals, int64_t* out_nrows, int64_t* out_nnz, void* stream) {
    SKFA_REQUIRE(indptr && keep && newid && out_indptr && out_nrows && out_nnz, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf<int32_t> cnt, flag;
    DevBuf<int64_t> start;
    if (int rc = cnt.alloc(nrows + 1)) return rc;
    if (int rc = flag.alloc(nrows + 1)) return rc;
    if (int rc = start.alloc(nrows + 1)) return rc;
    k_keep_flag<<<grid_of(nrows, 256), 256, 0, s>>>(keep, nrows, flag.p);
    SKFA_LAUNCH_CHECK();
    if (int rc = exclusive_scan_i32_to_i64(flag.p, newid, nrows, s)) return rc;
    k_keep_count<<<grid_of(nrows, 256), 256, 0, s>>>(indptr, indices, keep, nrows, cnt.p);
    SKFA_LAUNCH_CHECK();
    if (int rc = exclusive_scan_i32_to_i64(cnt.p, start.p, nrows, s)) return rc;
    SKFA_HIP(hipMemcpyAsync(out_nrows, newid + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipMemcpyAsync(out_nnz, start.p + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    k_condense_fill<<<grid_of(nrows, 256), 256, 0, s>>>(indptr, indices, vals, keep, newid, start.p, nrows, out_indptr,
                                                        out_indices, out_vals);
    SKFA_LAUNCH_CHECK();
    SKFA_HIP(hipMemcpyAsync(out_indptr + *out_nrows, out_nnz, sizeof(int64_t), hipMemcpyHostToDevice, s));
    SKFA_HIP(hipStreamSynchronize(s));
    return kOk;
}

int skfa_csr_lower(const int64_t* indptr, const int32_t* indices, int64_t nrows, int64_t* lptr, int32_t* lidx,
                   int64_t* lsrc, int64_t* nnz_lower, void* stream) {
    SKFA_REQUIRE(indptr && lptr && nnz_lower, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf<int32_t> cnt;
    if (int rc = cnt.alloc(nrows + 1)) return rc;
    k_lower_count<<<grid_of(nrows, 256), 256, 0, s>>>(indptr, indices, nrows, cnt.p);
    SKFA_LAUNCH_CHECK();
    if (int rc = exclusive_scan_i32_to_i64(cnt.p, lptr, nrows, s)) return rc;
    SKFA_HIP(hipMemcpyAsync(nnz_lower, lptr + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (lidx && lsrc) {
        k_lower_fill<<<grid_of(nrows, 256), 256, 0, s>>>(indptr, indices, nrows, lptr, lidx, lsrc);
        SKFA_LAUNCH_CHECK();
    }
    SKFA_HIP(hipStreamSynchronize(s));
    return kOk;
}

int skfa_csr_transpose(const int64_t* ptr, const int32_t* idx, int64_t nrows, int64_t ncols, int64_t* tptr,
                       int32_t* tidx, int64_t* tsrc, void* stream) {
    SKFA_REQUIRE(ptr && tptr && tidx && tsrc, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t nnz = 0;
    SKFA_HIP(hipMemcpyAsync(&nnz, ptr + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    DevBuf<int32_t> cnt;
    if (int rc = cnt.alloc(ncols + 1)) return rc;
    SKFA_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (ncols + 1), s));
    if (nnz) {
        k_col_count<<<grid_of(nnz, 256), 256, 0, s>>>(idx, nnz, cnt.p);
        SKFA_LAUNCH_CHECK();
    }
    if (int rc = exclusive_scan_i32_to_i64(cnt.p, tptr, ncols, s)) return rc;
    SKFA_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (ncols + 1), s));
    k_transpose_fill<<<grid_of(nrows, 256), 256, 0, s>>>(ptr, idx, nrows, tptr, cnt.p, tidx, tsrc);
    SKFA_LAUNCH_CHECK();
    k_transpose_sort<<<grid_of(ncols, 256), 256, 0, s>>>(tptr, tidx, tsrc, ncols);
    SKFA_LAUNCH_CHECK();
    SKFA_HIP(hipStreamSynchronize(s));
    return kOk;
}

int skfa_gather_values(const double* src, const int64_t* map, int64_t n, double* out, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) return kOk;
    k_gather_values<<<grid_of(n, 256), 256, 0, s>>>(src, map, n, out);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_gather_values_f32(const float* src, const int64_t* map, int64_t n, float* out, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) return kOk;
    k_gather_values<<<grid_of(n, 256), 256, 0, s>>>(src, map, n, out);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_diag_inverse(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t n, double* dinv,
                      void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) return kOk;
    k_diag_inv<<<grid_of(n, 256), 256, 0, s>>>(indptr, indices, vals, n, dinv);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_sys_dot(const double* x, const double* y, const int64_t* sys_ptr, int32_t nsys, double* out_xy,
                 double* out_xx, void* stream) {
    SKFA_REQUIRE(x && y && sys_ptr && out_xy, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nsys == 0) return kOk;
    k_sys_dot2<<<nsys, kRedThreads, 0, s>>>(x, y, sys_ptr, out_xy, out_xx);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

/* Batched PCG on a block-diagonal SPD system (all FP64).  Preconditioner:
 * kind 0 = Jacobi (dinv), kind 1 = factorised approximate inverse
 * z = G^T (G r) with G (lower CSR, FP32 values, applied in FP64) and
 * GT = G^T (CSR).  x is overwritten
 * (start from zero).  Stops each system at ||r||/||b|| <= tol (recursive
 * residual) or at max_iter; iters[s] receives the iteration counts.  The
 * convergence flags are read back every `check_every` iterations.
 * work: 6*n doubles; sys_work: 8*nsys doubles + 3*nsys+1 ints (device). */
int skfa_pcg_solve(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t n,
                   const int64_t* sys_ptr, int32_t nsys, const double* b, double* x, int32_t kind, const double* dinv,
                   const int64_t* gptr, const int32_t* gidx, const float* gval, const int64_t* gtptr,
                   const int32_t* gtidx, const float* gtval, double tol, int32_t max_iter, int32_t check_every,
                   int32_t spmv_lanes, double* work, double* sys_work, int32_t* int_work, int32_t* iters,
                   int32_t* total_iters, void* stream) {
    SKFA_REQUIRE(indptr && b && x && sys_ptr && work && sys_work && int_work && iters, "null argument");
    SKFA_REQUIRE(kind == 0 ? dinv != nullptr : (gptr && gidx && gval && gtptr && gtidx && gtval),
                 "preconditioner arrays are null");
    SKFA_REQUIRE(check_every >= 1, "check_every must be positive");
    hipStream_t s = static_cast<hipStream_t>(stream);
    double* r = work;
    double* z = work + n;
    double* p = work + 2 * n;
    double* q = work + 3 * n;
    double* t = work + 4 * n;
    double* rz = sys_work;
    double* rz_new = sys_work + nsys;
    double* pq = sys_work + 2 * nsys;
    double* rr = sys_work + 3 * nsys;
    double* bb = sys_work + 4 * nsys;
    int32_t* row_sys = reinterpret_cast<int32_t*>(work + 5 * n);
    int32_t* active = int_work;
    int32_t* nactive = int_work + nsys;
    const unsigned g = grid_of(n, 256);
    const double tol2 = tol * tol;
    auto precond = [&](const double* rin, double* zout) -> int {
        if (kind == 0) {
            k_jacobi<<<g, 256, 0, s>>>(rin, dinv, zout, n);
            SKFA_LAUNCH_CHECK();
            return kOk;
        }
        if (int rc = spmv_launch(gptr, gidx, gval, rin, t, n, 8, s)) return rc;
        return spmv_launch(gtptr, gtidx, gtval, t, zout, n, 8, s);
    };
    SKFA_REQUIRE(nsys <= 65535, "at most 65535 systems per batch");
    SKFA_HIP(hipMemsetAsync(iters, 0, sizeof(int32_t) * nsys, s));
    k_row_system<<<dim3(64, unsigned(nsys)), 256, 0, s>>>(sys_ptr, nsys, row_sys);
    SKFA_LAUNCH_CHECK();
    SKFA_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    SKFA_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    k_sys_dot2<<<nsys, kRedThreads, 0, s>>>(b, b, sys_ptr, bb, nullptr);
    // systems with b = 0 are solved by x = 0 and never iterate
    k_pcg_init<<<(nsys + 255) / 256, 256, 0, s>>>(bb, nsys, active);
    if (int rc = precond(r, z)) return rc;
    k_sys_dot2<<<nsys, kRedThreads, 0, s>>>(r, z, sys_ptr, rz, nullptr);
    k_pcg_direction<<<g, 256, 0, s>>>(p, z, rz, rz, active, row_sys, n, 1);
    SKFA_LAUNCH_CHECK();
    int it = 0;
    int32_t h_active = 1;
    while (it < max_iter && h_active > 0) {
        for (int k = 0; k < check_every && it < max_iter; ++k, ++it) {
            if (int rc = spmv_launch(indptr, indices, vals, p, q, n, spmv_lanes, s)) return rc;
            k_sys_dot2<<<nsys, kRedThreads, 0, s>>>(p, q, sys_ptr, pq, nullptr);
            k_pcg_update<<<g, 256, 0, s>>>(x, r, p, q, rz, pq, active, row_sys, n);
            if (int rc = precond(r, z)) return rc;
            k_sys_dot2<<<nsys, kRedThreads, 0, s>>>(r, z, sys_ptr, rz_new, rr);
            SKFA_HIP(hipMemsetAsync(nactive, 0, sizeof(int32_t), s));
            k_pcg_check<<<(nsys + 255) / 256, 256, 0, s>>>(rr, bb, tol2, nsys, active, iters, nactive);
            k_pcg_direction<<<g, 256, 0, s>>>(p, z, rz_new, rz, active, row_sys, n, 0);
            SKFA_LAUNCH_CHECK();
            std::swap(rz, rz_new);
        }
        SKFA_HIP(hipMemcpyAsync(&h_active, nactive, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SKFA_HIP(hipStreamSynchronize(s));
    }
    if (total_iters) *total_iters = it;
    return kOk;
}

}  // extern "C"
