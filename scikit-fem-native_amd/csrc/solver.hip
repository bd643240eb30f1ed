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
// Sixteen lanes per row (four rows per wave, blocks of 256 threads): lane
// `sub` of a group takes the row's entries base + sub in chunks of 16, so a
// row's index and value reads are coalesced; selected entries are compacted
// in order with the group's ballot.  The chunk loop bound is uniform over a
// group.
__device__ __forceinline__ uint32_t group_ballot(bool p) {
    return uint32_t(__ballot(p) >> (threadIdx.x & 48)) & 0xFFFFu;
}
__device__ __forceinline__ int lanes_before(uint32_t mask, int sub) {
    return __popc(mask & ((1u << sub) - 1u));
}

__global__ void __launch_bounds__(256) k_lower_count(const int64_t* __restrict__ indptr,
                                                     const int32_t* __restrict__ indices, int64_t nrows,
                                                     int32_t* __restrict__ count) {
    const int sub = threadIdx.x & 15;
    const int64_t groups = (int64_t(gridDim.x) * blockDim.x) >> 4;
    for (int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4; r < nrows; r += groups) {
        const int64_t b = indptr[r], len = indptr[r + 1] - b;
        int32_t c = 0;
        for (int64_t base = 0; base < len; base += 16) {
            const int64_t k = base + sub;
            c += __popc(group_ballot(k < len && indices[b + k] <= r));
        }
        if (sub == 0) count[r] = c;
    }
}

__global__ void __launch_bounds__(256) k_lower_fill(const int64_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices, int64_t nrows,
                                                    const int64_t* __restrict__ lptr, int32_t* __restrict__ lidx,
                                                    int64_t* __restrict__ lsrc) {
    const int sub = threadIdx.x & 15;
    const int64_t groups = (int64_t(gridDim.x) * blockDim.x) >> 4;
    for (int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4; r < nrows; r += groups) {
        const int64_t b = indptr[r], len = indptr[r + 1] - b;
        int64_t o = lptr[r];
        for (int64_t base = 0; base < len; base += 16) {
            const int64_t k = base + sub;
            const int32_t c = k < len ? indices[b + k] : 0;
            const bool take = k < len && c <= r;
            const uint32_t m = group_ballot(take);
            if (take) {
                const int64_t at = o + lanes_before(m, sub);
                lidx[at] = c;
                lsrc[at] = b + k;
            }
            o += __popc(m);
        }
    }
}

// Transpose of the lower pattern of a matrix whose pattern is symmetric:
// row i of L^T holds the columns j >= i of A's row i, and entry (i, j) of
// L^T sits in row j of L, found by binary search.  No atomics, no sort.
__global__ void __launch_bounds__(256) k_upper_count(const int64_t* __restrict__ indptr,
                                                     const int32_t* __restrict__ indices, int64_t nrows,
                                                     int32_t* __restrict__ count) {
    const int sub = threadIdx.x & 15;
    const int64_t groups = (int64_t(gridDim.x) * blockDim.x) >> 4;
    for (int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4; r < nrows; r += groups) {
        const int64_t b = indptr[r], len = indptr[r + 1] - b;
        int32_t c = 0;
        for (int64_t base = 0; base < len; base += 16) {
            const int64_t k = base + sub;
            c += __popc(group_ballot(k < len && indices[b + k] >= r));
        }
        if (sub == 0) count[r] = c;
    }
}

__global__ void __launch_bounds__(256) k_upper_fill(const int64_t* __restrict__ indptr,
                                                    const int32_t* __restrict__ indices, int64_t nrows,
                                                    const int64_t* __restrict__ lptr, const int32_t* __restrict__ lidx,
                                                    const int64_t* __restrict__ tptr, int32_t* __restrict__ tidx,
                                                    int64_t* __restrict__ tsrc, int32_t* __restrict__ missing) {
    const int sub = threadIdx.x & 15;
    const int64_t groups = (int64_t(gridDim.x) * blockDim.x) >> 4;
    for (int64_t r = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4; r < nrows; r += groups) {
        const int64_t b = indptr[r], len = indptr[r + 1] - b;
        int64_t o = tptr[r];
        for (int64_t base = 0; base < len; base += 16) {
            const int64_t k = base + sub;
            const int32_t j = k < len ? indices[b + k] : 0;
            const bool take = k < len && j >= r;
            const uint32_t m = group_ballot(take);
            if (take) {
                int64_t lo = lptr[j], hi = lptr[j + 1];
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (lidx[mid] < r) lo = mid + 1;
                    else hi = mid;
                }
                if (lo == lptr[j + 1] || lidx[lo] != r) {
                    atomicAdd(missing, 1);
                    lo = lptr[j];
                }
                const int64_t at = o + lanes_before(m, sub);
                tidx[at] = j;
                tsrc[at] = lo;
            }
            o += __popc(m);
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

__global__ void k_diag_inv(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                           const double* __restrict__ vals, int64_t n, double* __restrict__ dinv) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += int64_t(gridDim.x) * blockDim.x) {
        double d = 0.;
        for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k)
            if (indices[k] == r) d = vals[k];
        dinv[r] = 1. / d;
    }
}

// ---- fused batched PCG ----------------------------------------------------
// Rows are grouped in blocks of at most kPcgThreads consecutive rows that never
// straddle two systems (block table: system, first row, end row).  A sparse
// product in a block first forms the products of the block's contiguous
// nonzero range with coalesced loads into LDS, then each thread adds its
// row's products in order ("stream" layout, good for short FEM rows).  Each
// heavy kernel returns at once for blocks of converged systems and writes one
// partial sum per block; k_sys_reduce adds a system's partials in block
// order, so every result is reproducible bit for bit.
constexpr int kPcgThreads = 256;
constexpr int64_t kStreamCap = 6144;   // max nonzeros of one block (48 KiB of LDS)

struct BlockTable {
    const int32_t* sys;
    const int64_t* row0;
    const int64_t* row1;
};

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD runs a contiguous range of row blocks and the neighbouring entries a
// block gathers stay in that XCD's L2 (a bijection for any block count).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b & 7, idx = b >> 3;
    const int q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + idx;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
    // sum over the block, fixed order; result valid in thread 0
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.;
    if (threadIdx.x == 0)
        for (int k = 0; k < int(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}

// Column of nonzero k of a block whose first row is r0: int32 indices are
// absolute, int16 ones are offsets from r0 (the block-relative copy that
// k_block_relative writes for the preconditioner factors).
__device__ __forceinline__ int32_t column_of(int32_t j, int64_t) { return j; }
__device__ __forceinline__ int32_t column_of(int16_t d, int64_t r0) { return int32_t(r0) + int32_t(d); }

// Row sums of A x over rows [r0, r1) with x given by `xv(j)`; thread t gets row r0 + t.
template <typename I, typename V, typename XV>
__device__ __forceinline__ double stream_rows(const int64_t* __restrict__ ip, const I* __restrict__ ix,
                                              const V* __restrict__ av, int64_t r0, int64_t r1, double* prod,
                                              XV xv) {
    const int64_t k0 = ip[r0], k1 = ip[r1];
    const int64_t row = r0 + threadIdx.x;
    const bool has_row = row < r1;
    // the row bounds are loaded early so that they overlap the product phase
    const int64_t rb = has_row ? ip[row] : 0, re = has_row ? ip[row + 1] : 0;
    const int cnt = int(k1 - k0);
    const int nt = int(blockDim.x);
    // eight independent loads per thread in flight; the matrix is read once
    // per product, so its loads are non-temporal and leave the caches to the
    // vectors the products gather
    for (int base = int(threadIdx.x); base < cnt; base += 8 * nt) {
        int32_t j[8];
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = base + u * nt;
            j[u] = k < cnt ? column_of(__builtin_nontemporal_load(ix + k0 + k), r0) : 0;
            v[u] = k < cnt ? double(__builtin_nontemporal_load(av + k0 + k)) : 0.;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = base + u * nt;
            if (k < cnt) prod[k] = v[u] * xv(j[u]);
        }
    }
    __syncthreads();
    double acc = 0.;
    if (has_row)
        for (int64_t k = rb - k0; k < re - k0; ++k) acc += prod[k];
    return acc;
}

// p_new = z + beta p_old (beta = 0 on the first step), q = A p_new,
// partial[blk] = p_new . q
__global__ void __launch_bounds__(kPcgThreads) k_pcg_direction_spmv(
    BlockTable bt, const int64_t* __restrict__ ip, const int32_t* __restrict__ ix, const double* __restrict__ av,
    const double* __restrict__ z, const double* __restrict__ p_old, double* __restrict__ p_new,
    double* __restrict__ q, const double* __restrict__ rz, const double* __restrict__ rz_old,
    const int32_t* __restrict__ active, int first, double* __restrict__ partial) {
    extern __shared__ double prod[];
    __shared__ double sh[kPcgThreads / 64];
    const int blk = xcd_block(int(blockIdx.x), int(gridDim.x));
    const int s = bt.sys[blk];
    if (!active[s]) return;
    const double beta = first ? 0. : rz[s] / rz_old[s];
    const int64_t r0 = bt.row0[blk], r1 = bt.row1[blk];
    double acc;
    if (first) acc = stream_rows(ip, ix, av, r0, r1, prod, [&](int32_t j) { return z[j]; });
    else acc = stream_rows(ip, ix, av, r0, r1, prod, [&](int32_t j) { return fma(beta, p_old[j], z[j]); });
    const int64_t row = r0 + threadIdx.x;
    double contrib = 0.;
    if (row < r1) {
        const double pn = first ? z[row] : fma(beta, p_old[row], z[row]);
        p_new[row] = pn;
        q[row] = acc;
        contrib = pn * acc;
    }
    const double t = block_sum(contrib, sh);
    if (threadIdx.x == 0) partial[blk] = t;
}

// out[s] = sum of partial over the blocks of system s (active systems only)
__global__ void __launch_bounds__(256) k_sys_reduce(const int64_t* __restrict__ bptr, const double* __restrict__ p0,
                                                    const double* __restrict__ p1, const int32_t* __restrict__ active,
                                                    double* __restrict__ out0, double* __restrict__ out1) {
    __shared__ double s0[256], s1[256];
    const int s = blockIdx.x;
    if (!active[s]) return;
    double a = 0., b = 0.;
    for (int64_t k = bptr[s] + threadIdx.x; k < bptr[s + 1]; k += 256) {
        a += p0[k];
        if (p1) b += p1[k];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s0[threadIdx.x] += s0[threadIdx.x + off];
            s1[threadIdx.x] += s1[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out0[s] = s0[0];
        if (out1) out1[s] = s1[0];
    }
}

// Jacobi step: a = rz/pq; x += a p; r -= a q; z = dinv r;
// partials r.z and r.r.  One thread per row.
__global__ void __launch_bounds__(kPcgThreads) k_pcg_update_jacobi(
    BlockTable bt, double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
    const double* __restrict__ p, const double* __restrict__ q, const double* __restrict__ dinv,
    const double* __restrict__ rz, const double* __restrict__ pq, const int32_t* __restrict__ active,
    double* __restrict__ part_rz, double* __restrict__ part_rr) {
    __shared__ double sh0[kPcgThreads / 64], sh1[kPcgThreads / 64];
    const int blk = xcd_block(int(blockIdx.x), int(gridDim.x));
    const int s = bt.sys[blk];
    if (!active[s]) return;
    const double a = pq[s] != 0. ? rz[s] / pq[s] : 0.;
    const int64_t row = bt.row0[blk] + threadIdx.x;
    double c_rz = 0., c_rr = 0.;
    if (row < bt.row1[blk]) {
        x[row] = fma(a, p[row], x[row]);
        const double rn = fma(-a, q[row], r[row]);
        r[row] = rn;
        const double zn = rn * dinv[row];
        z[row] = zn;
        c_rz = rn * zn;
        c_rr = rn * rn;
    }
    const double t0 = block_sum(c_rz, sh0);
    const double t1 = block_sum(c_rr, sh1);
    if (threadIdx.x == 0) {
        part_rz[blk] = t0;
        part_rr[blk] = t1;
    }
}

// Factor step, first half: a = rz/pq; r_new = r_old - a q; x += a p;
// t = G r_new (neighbours' r_new formed on the fly from r_old and q);
// partial r.r.
template <typename I>
__global__ void __launch_bounds__(kPcgThreads) k_pcg_update_g(
    BlockTable bt, const int64_t* __restrict__ gp, const I* __restrict__ gi, const float* __restrict__ gv,
    double* __restrict__ x, const double* __restrict__ r_old, double* __restrict__ r_new, double* __restrict__ t,
    const double* __restrict__ p, const double* __restrict__ q, const double* __restrict__ rz,
    const double* __restrict__ pq, const int32_t* __restrict__ active, double* __restrict__ part_rr) {
    extern __shared__ double prod[];
    __shared__ double sh[kPcgThreads / 64];
    const int blk = xcd_block(int(blockIdx.x), int(gridDim.x));
    const int s = bt.sys[blk];
    if (!active[s]) return;
    const double a = pq[s] != 0. ? rz[s] / pq[s] : 0.;
    const int64_t r0 = bt.row0[blk], r1 = bt.row1[blk];
    const double acc = stream_rows(gp, gi, gv, r0, r1, prod, [&](int32_t j) { return fma(-a, q[j], r_old[j]); });
    const int64_t row = r0 + threadIdx.x;
    double c = 0.;
    if (row < r1) {
        const double rn = fma(-a, q[row], r_old[row]);
        r_new[row] = rn;
        x[row] = fma(a, p[row], x[row]);
        t[row] = acc;
        c = rn * rn;
    }
    const double tot = block_sum(c, sh);
    if (threadIdx.x == 0) part_rr[blk] = tot;
}

// Factor step, second half: z = G^T t; partial r.z.
template <typename I>
__global__ void __launch_bounds__(kPcgThreads) k_pcg_apply_gt(
    BlockTable bt, const int64_t* __restrict__ tp, const I* __restrict__ ti, const float* __restrict__ tv,
    const double* __restrict__ t, const double* __restrict__ r, double* __restrict__ z,
    const int32_t* __restrict__ active, double* __restrict__ part_rz) {
    extern __shared__ double prod[];
    __shared__ double sh[kPcgThreads / 64];
    const int blk = xcd_block(int(blockIdx.x), int(gridDim.x));
    const int s = bt.sys[blk];
    if (!active[s]) return;
    const int64_t r0 = bt.row0[blk], r1 = bt.row1[blk];
    const double acc = stream_rows(tp, ti, tv, r0, r1, prod, [&](int32_t j) { return t[j]; });
    const int64_t row = r0 + threadIdx.x;
    double c = 0.;
    if (row < r1) {
        z[row] = acc;
        c = r[row] * acc;
    }
    const double tot = block_sum(c, sh);
    if (threadIdx.x == 0) part_rz[blk] = tot;
}

// Largest nonzero count of one block of rows (for the LDS size).
__global__ void k_block_nnz_max(BlockTable bt, int64_t nblk, const int64_t* __restrict__ ip,
                                unsigned long long* __restrict__ out) {
    for (int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < nblk; b += int64_t(gridDim.x) * blockDim.x)
        atomicMax(out, static_cast<unsigned long long>(ip[bt.row1[b]] - ip[bt.row0[b]]));
}

// Block-relative 16-bit column offsets of a factor: out[k] = col - r0 for
// every nonzero k of block b (rows [r0, r1)); `wide` is set when an offset
// does not fit, and the caller keeps the int32 indices.  One workgroup per
// block (grid-stride over blocks).
__global__ void __launch_bounds__(256) k_block_relative(BlockTable bt, int64_t nblk, const int64_t* __restrict__ ip,
                                                        const int32_t* __restrict__ ix, int16_t* __restrict__ out,
                                                        int32_t* __restrict__ wide) {
    for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int64_t r0 = bt.row0[b];
        const int64_t k1 = ip[bt.row1[b]];
        for (int64_t k = ip[r0] + threadIdx.x; k < k1; k += blockDim.x) {
            const int64_t d = int64_t(ix[k]) - r0;
            if (d < -32768 || d > 32767) *wide = 1;
            out[k] = int16_t(d);
        }
    }
}

// After a step: rz_old <- rz, rz <- rz_new, iteration count, convergence.
__global__ void k_jacobi_init(const double* __restrict__ r, const double* __restrict__ dinv, double* __restrict__ z,
                              int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        z[i] = r[i] * dinv[i];
}

__global__ void k_pcg_advance(const double* __restrict__ rz_new, const double* __restrict__ rr,
                              const double* __restrict__ bb, double tol2, int32_t nsys, double* __restrict__ rz,
                              double* __restrict__ rz_old, int32_t* __restrict__ active, int32_t* __restrict__ iters,
                              int32_t* __restrict__ nactive) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys || !active[s]) return;
    iters[s] += 1;
    rz_old[s] = rz[s];
    rz[s] = rz_new[s];
    if (!(rr[s] > tol2 * bb[s])) active[s] = 0;
    else atomicAdd(nactive, 1);
}

__global__ void k_pcg_start(const double* __restrict__ bb, int32_t nsys, int32_t* __restrict__ active,
                            int32_t* __restrict__ iters) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    active[s] = bb[s] > 0. ? 1 : 0;
    iters[s] = 0;
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
    k_lower_count<<<grid_of(nrows * 16, 256), 256, 0, s>>>(indptr, indices, nrows, cnt.p);
    SKFA_LAUNCH_CHECK();
    if (int rc = exclusive_scan_i32_to_i64(cnt.p, lptr, nrows, s)) return rc;
    SKFA_HIP(hipMemcpyAsync(nnz_lower, lptr + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (lidx && lsrc) {
        k_lower_fill<<<grid_of(nrows * 16, 256), 256, 0, s>>>(indptr, indices, nrows, lptr, lidx, lsrc);
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

int skfa_csr_transpose_lower(const int64_t* indptr, const int32_t* indices, int64_t nrows, const int64_t* lptr,
                             const int32_t* lidx, int64_t* tptr, int32_t* tidx, int64_t* tsrc, void* stream) {
    SKFA_REQUIRE(indptr && indices && lptr && lidx && tptr && tidx && tsrc, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevBuf<int32_t> cnt, missing;
    if (int rc = cnt.alloc(size_t(nrows) + 1)) return rc;
    if (int rc = missing.alloc(1)) return rc;
    SKFA_HIP(hipMemsetAsync(missing.p, 0, sizeof(int32_t), s));
    k_upper_count<<<grid_of(nrows * 16, 256), 256, 0, s>>>(indptr, indices, nrows, cnt.p);
    SKFA_LAUNCH_CHECK();
    if (int rc = exclusive_scan_i32_to_i64(cnt.p, tptr, nrows, s)) return rc;
    k_upper_fill<<<grid_of(nrows * 16, 256), 256, 0, s>>>(indptr, indices, nrows, lptr, lidx, tptr, tidx, tsrc,
                                                        missing.p);
    SKFA_LAUNCH_CHECK();
    int32_t h_missing = 0;
    SKFA_HIP(hipMemcpyAsync(&h_missing, missing.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    SKFA_REQUIRE(h_missing == 0, "matrix pattern is not symmetric");
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

/* Batched PCG on a block-diagonal SPD system (all FP64 arithmetic).
 * kind 0 = Jacobi (dinv); kind 1 = factorised approximate inverse
 * z = G^T (G r) with G (lower CSR, FP32 values) and GT = G^T (CSR).
 * x is overwritten (x0 = 0).  System s stops when ||r|| <= tol ||b||
 * (recursive residual) or after max_iter; converged systems stop costing
 * bandwidth at once (their row blocks return early).  The host reads the
 * number of active systems every `check_every` iterations.
 * work: 7n doubles; sys_work: 8*nsys doubles; int_work: nsys+1 int32. */
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
    SKFA_REQUIRE(nsys >= 1, "at least one system");
    (void)spmv_lanes;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // block table: blocks of `rows_per_block` consecutive rows inside one
    // system, halved until every block's nonzeros fit the LDS staging area
    std::vector<int64_t> hsys(size_t(nsys) + 1);
    SKFA_HIP(hipMemcpyAsync(hsys.data(), sys_ptr, sizeof(int64_t) * (nsys + 1), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    SKFA_REQUIRE(hsys[0] == 0 && hsys[nsys] == n, "sys_ptr must start at 0 and end at n");
    for (int32_t q = 0; q < nsys; ++q) SKFA_REQUIRE(hsys[q + 1] >= hsys[q], "sys_ptr must be non-decreasing");
    std::vector<int32_t> bsys;
    std::vector<int64_t> br0, br1, bptr(size_t(nsys) + 1, 0);
    DevBuf<int32_t> d_bsys;
    DevBuf<int64_t> d_br0, d_br1, d_bptr;
    DevBuf<unsigned long long> d_max;
    if (int rc = d_max.alloc(3)) return rc;
    int64_t nblk = 0;
    int64_t max_block_nnz = 0;
    unsigned long long hmax[3] = {0, 0, 0};
    // half as many rows as threads: each thread forms about one row's worth
    // of products (one round of loads in flight) in the stream kernels
    for (int rows_per_block = kPcgThreads / 2;; rows_per_block /= 2) {
        bsys.clear();
        br0.clear();
        br1.clear();
        for (int32_t q = 0; q < nsys; ++q) {
            for (int64_t r = hsys[q]; r < hsys[q + 1]; r += rows_per_block) {
                bsys.push_back(q);
                br0.push_back(r);
                br1.push_back(std::min<int64_t>(r + rows_per_block, hsys[q + 1]));
            }
            bptr[q + 1] = int64_t(bsys.size());
        }
        nblk = int64_t(bsys.size());
        SKFA_REQUIRE(nblk < (int64_t(1) << 31), "too many row blocks");
        if (int rc = d_bsys.alloc(size_t(std::max<int64_t>(nblk, 1)))) return rc;
        if (int rc = d_br0.alloc(size_t(std::max<int64_t>(nblk, 1)))) return rc;
        if (int rc = d_br1.alloc(size_t(std::max<int64_t>(nblk, 1)))) return rc;
        if (nblk) {
            SKFA_HIP(hipMemcpyAsync(d_bsys.p, bsys.data(), sizeof(int32_t) * nblk, hipMemcpyHostToDevice, s));
            SKFA_HIP(hipMemcpyAsync(d_br0.p, br0.data(), sizeof(int64_t) * nblk, hipMemcpyHostToDevice, s));
            SKFA_HIP(hipMemcpyAsync(d_br1.p, br1.data(), sizeof(int64_t) * nblk, hipMemcpyHostToDevice, s));
        }
        SKFA_HIP(hipMemsetAsync(d_max.p, 0, 3 * sizeof(unsigned long long), s));
        const BlockTable probe{d_bsys.p, d_br0.p, d_br1.p};
        if (nblk) {
            k_block_nnz_max<<<grid_of(nblk, 256), 256, 0, s>>>(probe, nblk, indptr, d_max.p);
            if (kind != 0) {
                k_block_nnz_max<<<grid_of(nblk, 256), 256, 0, s>>>(probe, nblk, gptr, d_max.p + 1);
                k_block_nnz_max<<<grid_of(nblk, 256), 256, 0, s>>>(probe, nblk, gtptr, d_max.p + 2);
            }
            SKFA_LAUNCH_CHECK();
        }
        SKFA_HIP(hipMemcpyAsync(hmax, d_max.p, sizeof(hmax), hipMemcpyDeviceToHost, s));
        SKFA_HIP(hipStreamSynchronize(s));
        max_block_nnz = int64_t(std::max(hmax[0], std::max(hmax[1], hmax[2])));
        if (max_block_nnz <= kStreamCap) break;
        if (rows_per_block == 1) {
            set_error("a matrix row has more nonzeros than the PCG staging area holds");
            return kCapacity;
        }
    }
    // LDS staging per kernel, sized by its own matrix's largest block
    const size_t lds_bytes = size_t(std::max<unsigned long long>(hmax[0], 1)) * sizeof(double);
    int rows_per_block_used = 1;
    for (int64_t b = 0; b < nblk; ++b)
        rows_per_block_used = std::max<int>(rows_per_block_used, int(br1[size_t(b)] - br0[size_t(b)]));
    const int upd_threads = std::max(64, (rows_per_block_used + 63) / 64 * 64);
    const size_t lds_g = size_t(std::max<unsigned long long>(hmax[1], 1)) * sizeof(double);
    const size_t lds_gt = size_t(std::max<unsigned long long>(hmax[2], 1)) * sizeof(double);
    DevBuf<double> parts;
    if (int rc = d_bptr.alloc(size_t(nsys) + 1)) return rc;
    if (int rc = parts.alloc(size_t(std::max<int64_t>(nblk, 1)) * 3)) return rc;
    SKFA_HIP(hipMemcpyAsync(d_bptr.p, bptr.data(), sizeof(int64_t) * (nsys + 1), hipMemcpyHostToDevice, s));
    const BlockTable bt{d_bsys.p, d_br0.p, d_br1.p};
    // factor indices as 16-bit offsets from their block's first row when
    // every offset fits (FEM factors are banded): half the index bytes of
    // the two factor products per iteration, same arithmetic
    DevBuf<int16_t> g16, gt16;
    bool rel16 = false;
    if (kind != 0 && nblk > 0) {
        int64_t nnz_g = 0, nnz_gt = 0;
        SKFA_HIP(hipMemcpyAsync(&nnz_g, gptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        SKFA_HIP(hipMemcpyAsync(&nnz_gt, gtptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        SKFA_HIP(hipStreamSynchronize(s));
        DevBuf<int32_t> wide;
        if (int rc = g16.alloc(size_t(std::max<int64_t>(nnz_g, 1)))) return rc;
        if (int rc = gt16.alloc(size_t(std::max<int64_t>(nnz_gt, 1)))) return rc;
        if (int rc = wide.alloc(1)) return rc;
        SKFA_HIP(hipMemsetAsync(wide.p, 0, sizeof(int32_t), s));
        const unsigned cg = unsigned(std::min<int64_t>(nblk, 65536));
        k_block_relative<<<cg, 256, 0, s>>>(bt, nblk, gptr, gidx, g16.p, wide.p);
        k_block_relative<<<cg, 256, 0, s>>>(bt, nblk, gtptr, gtidx, gt16.p, wide.p);
        SKFA_LAUNCH_CHECK();
        int32_t hw = 1;
        SKFA_HIP(hipMemcpyAsync(&hw, wide.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SKFA_HIP(hipStreamSynchronize(s));
        rel16 = hw == 0;
    }
    double* part_a = parts.p;
    double* part_b = parts.p + nblk;
    double* part_c = parts.p + 2 * nblk;

    double* r0 = work;
    double* r1 = work + n;
    double* z = work + 2 * n;
    double* p0 = work + 3 * n;
    double* p1 = work + 4 * n;
    double* q = work + 5 * n;
    double* t = work + 6 * n;
    double* rz = sys_work;
    double* rz_old = sys_work + nsys;
    double* rz_new = sys_work + 2 * nsys;
    double* pq = sys_work + 3 * nsys;
    double* rr = sys_work + 4 * nsys;
    double* bb = sys_work + 5 * nsys;
    int32_t* active = int_work;
    int32_t* nactive = int_work + nsys;
    const unsigned g = grid_of(n, 256);
    const unsigned sg = unsigned((nsys + 255) / 256);
    const double tol2 = tol * tol;

    // start: x = 0, r = b, z = M r, rz = r.z, bb = b.b
    SKFA_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    SKFA_HIP(hipMemcpyAsync(r0, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    k_sys_dot2<<<nsys, kRedThreads, 0, s>>>(b, b, sys_ptr, bb, nullptr);
    k_pcg_start<<<sg, 256, 0, s>>>(bb, nsys, active, iters);
    if (kind == 0) {
        k_jacobi_init<<<g, 256, 0, s>>>(r0, dinv, z, n);
    } else {
        if (int rc = spmv_launch(gptr, gidx, gval, r0, t, n, 4, s)) return rc;
        if (int rc = spmv_launch(gtptr, gtidx, gtval, t, z, n, 4, s)) return rc;
    }
    k_sys_dot2<<<nsys, kRedThreads, 0, s>>>(r0, z, sys_ptr, rz, nullptr);
    SKFA_LAUNCH_CHECK();
    int it = 0;
    int32_t h_active = 1;
    double* p_old = p1;
    double* p_new = p0;
    double* r_cur = r0;
    double* r_nxt = r1;
    const unsigned nb = unsigned(std::max<int64_t>(nblk, 1));
    while (it < max_iter && h_active > 0) {
        SKFA_HIP(hipMemsetAsync(nactive, 0, sizeof(int32_t), s));
        for (int k = 0; k < check_every && it < max_iter; ++k, ++it) {
            k_pcg_direction_spmv<<<nb, kPcgThreads, lds_bytes, s>>>(
                bt, indptr, indices, vals, z, p_old, p_new, q, rz, rz_old, active, it == 0 ? 1 : 0, part_a);
            k_sys_reduce<<<nsys, 256, 0, s>>>(d_bptr.p, part_a, nullptr, active, pq, nullptr);
            if (kind == 0) {
                k_pcg_update_jacobi<<<nb, upd_threads, 0, s>>>(bt, x, r_cur, z, p_new, q, dinv, rz, pq, active,
                                                            part_b, part_c);
            } else {
                if (rel16) {
                    k_pcg_update_g<int16_t><<<nb, kPcgThreads, lds_g, s>>>(bt, gptr, g16.p, gval, x, r_cur, r_nxt,
                                                                          t, p_new, q, rz, pq, active, part_c);
                    k_pcg_apply_gt<int16_t><<<nb, kPcgThreads, lds_gt, s>>>(bt, gtptr, gt16.p, gtval, t, r_nxt, z,
                                                                           active, part_b);
                } else {
                    k_pcg_update_g<int32_t><<<nb, kPcgThreads, lds_g, s>>>(bt, gptr, gidx, gval, x, r_cur, r_nxt,
                                                                          t, p_new, q, rz, pq, active, part_c);
                    k_pcg_apply_gt<int32_t><<<nb, kPcgThreads, lds_gt, s>>>(bt, gtptr, gtidx, gtval, t, r_nxt, z,
                                                                           active, part_b);
                }
                std::swap(r_cur, r_nxt);
            }
            k_sys_reduce<<<nsys, 256, 0, s>>>(d_bptr.p, part_b, part_c, active, rz_new, rr);
            if (k == check_every - 1 || it == max_iter - 1)
                SKFA_HIP(hipMemsetAsync(nactive, 0, sizeof(int32_t), s));
            k_pcg_advance<<<sg, 256, 0, s>>>(rz_new, rr, bb, tol2, nsys, rz, rz_old, active, iters, nactive);
            SKFA_LAUNCH_CHECK();
            std::swap(p_old, p_new);
        }
        SKFA_HIP(hipMemcpyAsync(&h_active, nactive, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SKFA_HIP(hipStreamSynchronize(s));
    }
    if (total_iters) *total_iters = it;
    return kOk;
}

}  // extern "C"
