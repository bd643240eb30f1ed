// Learned factorised sparse approximate inverse (FSAI) for batched SPD
// systems: per-nonzero features, the Kaporin objective and its gradient, a
// 16-128-128-1 ReLU MLP on BF16 MFMA (FP32 accumulate) for inference and
// training, and an Adam step.  No counterpart in the reference (its solver
// policy is external, python/skfemntv/capabilities.py:67).
//
// Notation.  A is SPD, block-diagonal over systems; D = diag(A);
// At = D^{-1/2} A D^{-1/2} has unit diagonal.  The preconditioner factor G
// lives on the lower-triangular pattern of A: G = Gt D^{-1/2}, where the
// network predicts Gt entry by entry (off-diagonal Gt_ij = o_ij, diagonal
// Gt_ii = exp(o_ii) > 0), so M^{-1} = G^T G is SPD by construction and a
// network that outputs zeros is exactly Jacobi.  The training objective per
// system (Kaporin's condition number of Gt At Gt^T, in logarithmic form) is
//   L = log(tr(Gt At Gt^T) / n) - (2/n) sum_i log Gt_ii,
// which is invariant to scaling Gt and minimised by the classical FSAI.
//
// MLP layout (FP32 master parameters, one flat vector):
//   W1[16][128] (in, out) | b1[128] | W2[128][128] (in, out) | b2[128] |
//   w3[128] | b3
// MFMA operands are BF16 copies of W1/W2 pre-arranged per lane ("fragments").
// Activations are kept transposed (hidden unit = MFMA row, nonzero = MFMA
// column), so one layer's accumulator layout is the next layer's B operand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "skfa_internal.h"
#include "../../include/skfem_amd.h"

namespace skfa {
namespace {

constexpr int kF = 16;    // input features
constexpr int kH = 128;   // hidden width
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kF * kH;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kParams = kOffB3 + 1;
static_assert(kParams == SKFA_FSAI_PARAMS, "parameter count");
constexpr int kFragW1 = 8 * 64 * 4;        // bf16 elements
constexpr int kFragW2 = 8 * 4 * 64 * 8;    // bf16 elements (forward or backward arrangement)
constexpr int kFrag = kFragW1 + 2 * kFragW2;
static_assert(kFrag == SKFA_FSAI_FRAG, "fragment count");
constexpr int kBiasGrads = 3 * kH + 1;     // b1, b2, w3, b3 gradient partials per wave

typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even (gfx950 v_cvt_pk_bf16_f32)
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
__host__ __device__ __forceinline__ int unit_of_slot(int kb2, int quad, int s) {
    // k-slot s (0..7) of lane quad `quad` in a 32-deep MFMA step kb2 maps to a
    // hidden unit so that two 16-row accumulator tiles feed it directly
    return kb2 * 32 + (s < 4 ? quad * 4 + s : 16 + quad * 4 + (s - 4));
}

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma32(v8bf a, v8bf b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v8bf as_v8bf(uint4 u) { return __builtin_bit_cast(v8bf, u); }
__device__ __forceinline__ v4s as_v4s(uint2 u) { return __builtin_bit_cast(v4s, u); }

// ------------------------------------------------------------------ features
// Per row: [0] sum_{k!=i} |At_ik|, [1] (sum_k A_ik)/A_ii, [2] row length/16,
// [3] mean distance to off-diagonal neighbours (1 if none).  16 lanes per
// row: lane l takes the row's nonzeros l, l+16, ... (coalesced), then the 16
// partial sums are combined by shuffles.
__device__ __forceinline__ double sum16(double v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

__global__ void __launch_bounds__(256) k_row_stats(const int64_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                                   const double* __restrict__ av, const double* __restrict__ rs,
                                                   const double* __restrict__ xyz, int64_t n, double* __restrict__ stats) {
    const int sub = threadIdx.x & 15;
    const int64_t groups = (int64_t(gridDim.x) * blockDim.x) >> 4;
    // the loop bound is uniform over each 16-lane group, which the shuffles need
    for (int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4; i < n; i += groups) {
        double abs_sum = 0., sum = 0., dist = 0., aii = 0., cnt = 0.;
        const double rsi = rs[i];
        double xi = 0., yi = 0., zi = 0.;
        if (xyz) {
            xi = xyz[3 * i];
            yi = xyz[3 * i + 1];
            zi = xyz[3 * i + 2];
        }
        for (int64_t k = ip[i] + sub; k < ip[i + 1]; k += 16) {
            const int32_t j = ix[k];
            const double a = av[k];
            sum += a;
            if (j == i) {
                aii = a;
                continue;
            }
            abs_sum += fabs(a * rsi * rs[j]);
            if (xyz) {
                const double dx = xyz[3 * j] - xi, dy = xyz[3 * j + 1] - yi, dz = xyz[3 * j + 2] - zi;
                dist += sqrt(dx * dx + dy * dy + dz * dz);
            }
            cnt += 1.;
        }
        abs_sum = sum16(abs_sum);
        sum = sum16(sum);
        dist = sum16(dist);
        cnt = sum16(cnt);
        aii = sum16(aii);   // one lane holds the diagonal, the others 0
        if (sub == 0) {
            stats[4 * i + 0] = abs_sum;
            stats[4 * i + 1] = sum / aii;
            stats[4 * i + 2] = double(ip[i + 1] - ip[i]) / 16.;
            stats[4 * i + 3] = (xyz && cnt > 0.) ? dist / cnt : 1.;
        }
    }
}

// 1/sqrt(A_ii), one thread per row.  A's rows are sorted, so the lower
// pattern's count of row i says where the diagonal sits: the last of the
// row's entries j <= i.  A row without a diagonal gives 1/sqrt(0).
__global__ void __launch_bounds__(256) k_rsqrt_diag(const int64_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                                    const double* __restrict__ av, const int64_t* __restrict__ lptr,
                                                    int64_t n, double* __restrict__ rs) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t k = ip[i] + (lptr[i + 1] - lptr[i]) - 1;
        const double d = (k >= ip[i] && ix[k] == i) ? av[k] : 0.;
        rs[i] = 1. / sqrt(d);
    }
}

// One thread per lower nonzero (i, j <= i).  Features (float):
//  0 At_ij            1 [i == j]          2 abs-sum(i)       3 abs-sum(j)
//  4 rowsum(i)/A_ii   5 rowsum(j)/A_jj    6 len(i)/16        7 len(j)/16
//  8 sum_{j<k<i} At_ik At_kj   9 sum_{k<j} At_ik At_jk   10 sum_k At_ik At_kj
// 11 log(A_ii/A_jj)/2 12-14 (x_j - x_i)/h_i  15 |x_j - x_i|/h_i
__device__ __forceinline__ int64_t row_of(const int64_t* __restrict__ lptr, int64_t n, int64_t q) {
    // largest i with lptr[i] <= q
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (lptr[mid] <= q) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Blocks take 256 consecutive lower nonzeros.  Every row holds at least its
// diagonal, so they span at most 256 rows: one search finds the first, and
// each thread then searches the block's window of row pointers in LDS.
// Rows i of the block are consecutive, so their A entries form one
// contiguous range, staged in LDS with coalesced loads (when it fits
// kFeatStage entries; else row i is read from global memory).  Each thread
// then walks row j of A in order, kFeatChunk entries at a time with their loads
// (column, value) issued together, and
// advances its pointer into row i as in a sorted merge; the matched columns
// come in increasing order, so every sum has the merge's order.
constexpr int kFeatBlock = 256;
constexpr int kFeatStage = 2048;
constexpr int kFeatChunk = 4;   // row-j entries whose loads are issued together

// val_i(a) = A_ik rs_k^2 for column k of row i, so the product of the two
// scaled entries At_ik At_jk = val_i(a) A_jk rs_i rs_j needs no gather of
// rs_k while row j is walked; At_ij = val_i(a) rs_i / rs_j at k = j.
template <typename ColI, typename ValI>
__device__ __forceinline__ void feature_merge(int64_t a, int64_t ea, ColI col_i, ValI val_i,
                                              const int32_t* __restrict__ ix, const double* __restrict__ av,
                                              int64_t bj, int64_t ej, int64_t i, int32_t j, double rsi, double rsj,
                                              double& aij, double& t_mid, double& t_lo, double& t_all) {
    const double rij = rsi * rsj;
    for (int64_t b0 = bj; b0 < ej; b0 += kFeatChunk) {
        int32_t d[kFeatChunk];
        double v[kFeatChunk];
#pragma unroll
        for (int u = 0; u < kFeatChunk; ++u) {
            const bool ok = b0 + u < ej;
            d[u] = ok ? ix[b0 + u] : 0x7fffffff;
            v[u] = ok ? av[b0 + u] : 0.;
        }
#pragma unroll
        for (int u = 0; u < kFeatChunk; ++u) {
            // (a padding entry, column 0x7fffffff, runs a to the end and matches nothing)
            while (a < ea && col_i(a) < d[u]) ++a;
            if (a < ea && col_i(a) == d[u]) {
                const double va = val_i(a);
                const double p = (va * v[u]) * rij;
                // (adds of 0 rather than branches: a branch between the two
                // sums puts them in scratch memory)
                t_all += p;
                t_lo += d[u] < j ? p : 0.;
                t_mid += (d[u] > j && d[u] < i) ? p : 0.;
                if (d[u] == j) aij = va * rsi / rsj;
                ++a;
            }
        }
    }
}

__global__ void __launch_bounds__(kFeatBlock) k_features(const int64_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                                         const double* __restrict__ av, const double* __restrict__ rs,
                                                         const double* __restrict__ stats, const double* __restrict__ xyz,
                                                         const int64_t* __restrict__ lptr, const int32_t* __restrict__ lidx,
                                                         int64_t n, int64_t nnz_l, float* __restrict__ feat) {
    __shared__ int64_t win[kFeatBlock + 1];
    __shared__ int64_t ipw[kFeatBlock + 1];
    __shared__ int32_t sc[kFeatStage];
    __shared__ double sv[kFeatStage];
    __shared__ int64_t first_row;
    __shared__ int last_lo;
    for (int64_t q0 = int64_t(blockIdx.x) * kFeatBlock; q0 < nnz_l; q0 += int64_t(gridDim.x) * kFeatBlock) {
        if (threadIdx.x == 0) first_row = row_of(lptr, n, q0);
        __syncthreads();
        const int64_t i0 = first_row;
        // win[k] = lptr[i0 + k]; rows past n read as nnz_l
        for (int k = threadIdx.x; k <= kFeatBlock; k += blockDim.x) {
            win[k] = i0 + k <= n ? lptr[i0 + k] : nnz_l;
            ipw[k] = i0 + k <= n ? ip[i0 + k] : ip[n];
        }
        __syncthreads();
        const int64_t q = q0 + threadIdx.x;
        int lo = 0, hi = kFeatBlock;   // largest k with win[k] <= q
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (win[mid] <= q) lo = mid;
            else hi = mid;
        }
        // the block's last nonzero (thread 255, or the last of the matrix) has the largest row
        if (q == min(q0 + kFeatBlock, nnz_l) - 1) last_lo = lo;
        __syncthreads();
        const int64_t base = ipw[0];
        const int64_t count = ipw[last_lo + 1] - base;
        const bool staged = count <= kFeatStage;
        if (staged)
            for (int64_t e = threadIdx.x; e < count; e += blockDim.x) {
                const int32_t c = ix[base + e];
                const double r = rs[c];
                sc[e] = c;
                sv[e] = av[base + e] * (r * r);
            }
        __syncthreads();
        if (q < nnz_l) {
            const int64_t i = i0 + lo;
            const int32_t j = lidx[q];
            const int64_t bj = ip[j], ej = ip[j + 1];
            const double rsi = rs[i], rsj = rs[j];
            double aij = 0., t_mid = 0., t_lo = 0., t_all = 0.;
            if (staged)
                feature_merge(ipw[lo] - base, ipw[lo + 1] - base, [&](int64_t e) { return sc[e]; },
                              [&](int64_t e) { return sv[e]; }, ix, av, bj, ej, i, j, rsi, rsj, aij, t_mid, t_lo,
                              t_all);
            else
                feature_merge(
                    ipw[lo], ipw[lo + 1], [&](int64_t e) { return ix[e]; },
                    [&](int64_t e) {
                        const double r = rs[ix[e]];
                        return av[e] * (r * r);
                    },
                    ix, av, bj, ej, i, j, rsi, rsj, aij, t_mid, t_lo, t_all);
            float f[kF];
            f[0] = float(aij);
            f[1] = (i == j) ? 1.f : 0.f;
            f[2] = float(stats[4 * i + 0]);
            f[3] = float(stats[4 * j + 0]);
            f[4] = float(stats[4 * i + 1]);
            f[5] = float(stats[4 * j + 1]);
            f[6] = float(stats[4 * i + 2]);
            f[7] = float(stats[4 * j + 2]);
            f[8] = float(t_mid);
            f[9] = float(t_lo);
            f[10] = float(t_all);
            f[11] = float(0.5 * log(rsj * rsj / (rsi * rsi)));
            if (xyz && i != j) {
                const double h = stats[4 * i + 3];
                const double dx = (xyz[3 * j] - xyz[3 * i]) / h, dy = (xyz[3 * j + 1] - xyz[3 * i + 1]) / h,
                             dz = (xyz[3 * j + 2] - xyz[3 * i + 2]) / h;
                f[12] = float(dx);
                f[13] = float(dy);
                f[14] = float(dz);
                f[15] = float(sqrt(dx * dx + dy * dy + dz * dz));
            } else {
                f[12] = f[13] = f[14] = f[15] = 0.f;
            }
            float4* out = reinterpret_cast<float4*>(feat + q * kF);
#pragma unroll
            for (int v = 0; v < 4; ++v) out[v] = make_float4(f[4 * v], f[4 * v + 1], f[4 * v + 2], f[4 * v + 3]);
        }
        __syncthreads();   // the staged rows and windows are rewritten by the next block step
    }
}

// ------------------------------------------------------------------ Kaporin objective
// Gt from network outputs; G = Gt D^{-1/2} as float; per-nonzero diag flag.
// Sixteen lanes per row, each taking the row's lower entries sub, sub+16, ...
__global__ void __launch_bounds__(256) k_gtilde(const float* __restrict__ o, const int64_t* __restrict__ lptr,
                                                const int32_t* __restrict__ lidx, const double* __restrict__ rs,
                                                int64_t n, double* __restrict__ gt, float* __restrict__ g,
                                                double* __restrict__ logdiag) {
    const int sub = threadIdx.x & 15;
    const int64_t groups = (int64_t(gridDim.x) * blockDim.x) >> 4;
    for (int64_t i = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 4; i < n; i += groups) {
        const int64_t b = lptr[i], e = lptr[i + 1];
        for (int64_t q = b + sub; q < e; q += 16) {
            const int32_t j = lidx[q];
            double v = double(o[q]);
            if (j == i) {
                v = fmin(fmax(v, -30.), 30.);
                if (logdiag) logdiag[i] = v;
                v = exp(v);
            }
            if (gt) gt[q] = v;
            if (g) g[q] = float(v * rs[j]);
        }
    }
}

// P_q = (Gt At)_ij for lower nonzero q = (i, j); prod_q = Gt_ij P_q.
__global__ void k_kaporin_p(const int64_t* __restrict__ ip, const int32_t* __restrict__ ix,
                            const double* __restrict__ av, const double* __restrict__ rs,
                            const int64_t* __restrict__ lptr, const int32_t* __restrict__ lidx,
                            const double* __restrict__ gt, int64_t n, double* __restrict__ P,
                            double* __restrict__ prod) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        for (int64_t q = lptr[i]; q < lptr[i + 1]; ++q) {
            const int32_t j = lidx[q];
            double acc = 0.;
            for (int64_t m = lptr[i]; m < lptr[i + 1]; ++m) {
                const int32_t k = lidx[m];
                // At_kj by binary search in row k
                int64_t lo = ip[k], hi = ip[k + 1];
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (ix[mid] < j) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < ip[k + 1] && ix[lo] == j) acc = fma(gt[m], av[lo] * rs[k] * rs[j], acc);
            }
            P[q] = acc;
            prod[q] = gt[q] * acc;
        }
    }
}

constexpr int kRed = 256;
// out[s] = sum of v over [seg[s], seg[s+1]) (fixed order)
__global__ void __launch_bounds__(kRed) k_segsum(const double* __restrict__ v, const int64_t* __restrict__ seg,
                                                 double* __restrict__ out) {
    __shared__ double sh[kRed];
    const int s = blockIdx.x;
    double a = 0.;
    for (int64_t k = seg[s] + threadIdx.x; k < seg[s + 1]; k += kRed) a += v[k];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int off = kRed / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = sh[0];
}

// Loss per system and gradient wrt network outputs o (scaled by `scale`):
//  dL/do_ij = 2 P_ij / tr                      (i != j)
//  dL/do_ii = 2 P_ii Gt_ii / tr - 2 / n_s      (Gt_ii = exp(o_ii))
__global__ void k_kaporin_loss(const double* __restrict__ tr, const double* __restrict__ logsum,
                               const int64_t* __restrict__ sys_ptr, int32_t nsys, double* __restrict__ loss) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const double ns = double(sys_ptr[s + 1] - sys_ptr[s]);
    loss[s] = log(tr[s] / ns) - 2. * logsum[s] / ns;
}

__global__ void k_kaporin_grad(const int64_t* __restrict__ lptr, const int32_t* __restrict__ lidx,
                               const double* __restrict__ gt, const double* __restrict__ P,
                               const double* __restrict__ tr, const int32_t* __restrict__ row_sys,
                               const int64_t* __restrict__ sys_ptr, int64_t n, double scale,
                               float* __restrict__ go) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const int s = row_sys[i];
        const double t = tr[s];
        const double ns = double(sys_ptr[s + 1] - sys_ptr[s]);
        for (int64_t q = lptr[i]; q < lptr[i + 1]; ++q) {
            double gr = 2. * P[q] / t;
            if (lidx[q] == i) gr = gr * gt[q] - 2. / ns;
            go[q] = float(gr * scale);
        }
    }
}

__global__ void k_row_sys(const int64_t* __restrict__ sys_ptr, int32_t nsys, int32_t* __restrict__ row_sys) {
    const int s = blockIdx.y;
    for (int64_t i = sys_ptr[s] + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < sys_ptr[s + 1];
         i += int64_t(gridDim.x) * blockDim.x)
        row_sys[i] = s;
}

__global__ void k_seg_of(const int64_t* __restrict__ sys_ptr, const int64_t* __restrict__ lptr, int32_t nsys,
                         int64_t* __restrict__ seg) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= nsys) seg[s] = lptr[sys_ptr[s]];
}

// ------------------------------------------------------------------ MLP
struct MlpArgs {
    const float* params;    // FP32 masters
    const uint16_t* frag;   // BF16 fragments: W1 | W2 forward | W2 backward
};

// Load one 16-row tile of features for this lane: rows tile0 + lane%16,
// features 4*(lane/16) .. +3, rounded to BF16.
__device__ __forceinline__ float4 load_x_raw(const float* __restrict__ feat, int64_t row, int64_t nnz, int quad) {
    return row < nnz ? *reinterpret_cast<const float4*>(feat + row * kF + quad * 4) : make_float4(0.f, 0.f, 0.f, 0.f);
}
__device__ __forceinline__ v4s x_bf16(float4 v) {
    uint2 u;
    u.x = uint32_t(f2bf(v.x)) | (uint32_t(f2bf(v.y)) << 16);
    u.y = uint32_t(f2bf(v.z)) | (uint32_t(f2bf(v.w)) << 16);
    return as_v4s(u);
}

// max(v, 0) as a signed-integer max of the bits: one instruction, where
// fmaxf first quiets a possible signalling NaN of an MFMA result.  Equal to
// fmaxf(v, 0) for every non-NaN v (-0 gives +0).
__device__ __forceinline__ float relu(float v) {
    return __builtin_bit_cast(float, max(__builtin_bit_cast(int32_t, v), 0));
}

typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
// Two BF16 values in one 32-bit word, `lo` in the low half (one packed
// round-to-nearest-even conversion).
typedef float v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
    const v2f f = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, v2bf));
}

typedef short v2s __attribute__((ext_vector_type(2)));

// Eight values as BF16 (round to nearest even); with `act`, ReLU is applied
// to the packed BF16 pairs as a signed 16-bit max with 0 (sign bit set =
// negative), which equals converting relu(v).
__device__ __forceinline__ uint4 pack8(const v4f& a, const v4f& b, bool act) {
    uint4 u = make_uint4(pack2(a[0], a[1]), pack2(a[2], a[3]), pack2(b[0], b[1]), pack2(b[2], b[3]));
    if (act) {
        const v2s z = {0, 0};
        u.x = __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(v2s, u.x), z));
        u.y = __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(v2s, u.y), z));
        u.z = __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(v2s, u.z), z));
        u.w = __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(v2s, u.w), z));
    }
    return u;
}


// Storage order of hidden units (and of input features) in the backward
// activation buffers: inside each block of 16 the two 2-bit fields are
// swapped (u = 4q + r -> 4r + q), so that the four lane groups q of one
// store instruction write four adjacent units (256 contiguous bytes).  The
// map is its own inverse.
__host__ __device__ __forceinline__ int unit_store(int u) {
    return (u & ~15) | ((u & 3) << 2) | ((u >> 2) & 3);
}

// Whether BF16 slot s (0..7) of a packed uint4 is nonzero.
__device__ __forceinline__ bool slot_nonzero(const uint4& u, int s) {
    const uint32_t w = s < 2 ? u.x : s < 4 ? u.y : s < 6 ? u.z : u.w;
    return (s & 1) ? w > 0xffffu : (w & 0xffffu) != 0u;
}

__device__ __forceinline__ float comp(const float4& v, int r) {
    return r == 0 ? v.x : r == 1 ? v.y : r == 2 ? v.z : v.w;
}

// Sum over the four lanes that share a column (lanes c, c+16, c+32, c+48).
__device__ __forceinline__ float col_sum(float v) {
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}
// Sum over the sixteen lanes that share a quad (same hidden units).
__device__ __forceinline__ float quad_sum(float v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

constexpr int kMlpWaves = 4;

// Stage the BF16 fragments (W1 and one W2 arrangement) in LDS.
__device__ __forceinline__ void stage_frags(const uint16_t* __restrict__ frag, int w2_offset, uint2* w1s, uint4* w2s) {
    const uint2* g1 = reinterpret_cast<const uint2*>(frag);
    const uint4* g2 = reinterpret_cast<const uint4*>(frag + w2_offset);
    for (int k = threadIdx.x; k < kFragW1 / 4; k += blockDim.x) w1s[k] = g1[k];
    for (int k = threadIdx.x; k < kFragW2 / 8; k += blockDim.x) w2s[k] = g2[k];
    __syncthreads();
}

// Inference of T 16-row tiles at once: each BF16 weight fragment read from
// LDS feeds T MFMAs (the fragments are the LDS traffic of this kernel).  The
// output unit's dot product is taken right after each 16-unit block of
// layer 2 (same summation order as summing the full z2 afterwards), so only
// one accumulator tile per input tile is live.  b1, b2 and w3 come from LDS
// (float4 per lane: units ob*16 + quad*4 + 0..3); the biases are the MFMAs'
// initial accumulators, which keeps the VALU work per tile below the MFMA
// time.  Per tile the arithmetic and its order do not depend on T.
template <int T>
__device__ __forceinline__ void tiles_forward_out(const uint2* __restrict__ w1f, const uint4* __restrict__ w2f,
                                                  const float4* __restrict__ b1s, const float4* __restrict__ b2s,
                                                  const float4* __restrict__ w3s, const v4s (&x)[T], int lane,
                                                  float (&o)[T]) {
    const int quad = lane >> 4;
    uint4 h1b[T][4];
#pragma unroll
    for (int kb2 = 0; kb2 < 4; ++kb2) {
        v4f z[T][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ob = 2 * kb2 + h;
            const v4s w = as_v4s(w1f[ob * 64 + lane]);
            const float4 b = b1s[ob * 4 + quad];
            const v4f bias = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int t = 0; t < T; ++t) z[t][h] = mfma16(w, x[t], bias);
        }
#pragma unroll
        for (int t = 0; t < T; ++t) h1b[t][kb2] = pack8(z[t][0], z[t][1], true);
    }
#pragma unroll
    for (int t = 0; t < T; ++t) o[t] = 0.f;
#pragma unroll
    for (int ob = 0; ob < 8; ++ob) {
        const float4 b = b2s[ob * 4 + quad];
        const v4f bias = {b.x, b.y, b.z, b.w};
        v4f c[T];
#pragma unroll
        for (int kb2 = 0; kb2 < 4; ++kb2) {
            const v8bf w = as_v8bf(w2f[(ob * 4 + kb2) * 64 + lane]);
            // (the first step reads the bias as its accumulator, no copies)
#pragma unroll
            for (int t = 0; t < T; ++t) c[t] = mfma32(w, as_v8bf(h1b[t][kb2]), kb2 == 0 ? bias : c[t]);
        }
        const float4 w = w3s[ob * 4 + quad];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            o[t] = fmaf(w.x, relu(c[t][0]), o[t]);
            o[t] = fmaf(w.y, relu(c[t][1]), o[t]);
            o[t] = fmaf(w.z, relu(c[t][2]), o[t]);
            o[t] = fmaf(w.w, relu(c[t][3]), o[t]);
        }
        // keep one block's fragments and accumulators live at a time
        __builtin_amdgcn_sched_barrier(0);
    }
}

constexpr int kFwdTiles = 2;   // 16-row tiles per pass of the inference kernel

__global__ void __launch_bounds__(64 * kMlpWaves) k_mlp_forward(MlpArgs a, const float* __restrict__ feat,
                                                                int64_t nnz, float* __restrict__ out) {
    constexpr int T = kFwdTiles;
    __shared__ uint2 w1s[kFragW1 / 4];
    __shared__ uint4 w2s[kFragW2 / 8];
    __shared__ float4 b1s[kH / 4], b2s[kH / 4], w3s[kH / 4];
    for (int k = threadIdx.x; k < kH; k += blockDim.x) {
        reinterpret_cast<float*>(b1s)[k] = a.params[kOffB1 + k];
        reinterpret_cast<float*>(b2s)[k] = a.params[kOffB2 + k];
        reinterpret_cast<float*>(w3s)[k] = a.params[kOffW3 + k];
    }
    stage_frags(a.frag, kFragW1, w1s, w2s);
    const int lane = threadIdx.x & 63;
    const int quad = lane >> 4;
    const int64_t groups = (nnz + 16 * T - 1) / (16 * T);
    const int64_t wave = int64_t(blockIdx.x) * kMlpWaves + (threadIdx.x >> 6);
    const int64_t nwaves = int64_t(gridDim.x) * kMlpWaves;
    const float b3 = a.params[kOffB3];
    // the next group's features are in flight while this group is computed
    float4 xn[T];
#pragma unroll
    for (int t = 0; t < T; ++t) xn[t] = load_x_raw(feat, (wave * T + t) * 16 + (lane & 15), nnz, quad);
    for (int64_t grp = wave; grp < groups; grp += nwaves) {
        v4s x[T];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            x[t] = x_bf16(xn[t]);
            xn[t] = load_x_raw(feat, ((grp + nwaves) * T + t) * 16 + (lane & 15), nnz, quad);
        }
        float o[T];
        tiles_forward_out<T>(w1s, w2s, b1s, b2s, w3s, x, lane, o);
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int64_t row = (grp * T + t) * 16 + (lane & 15);
            const float v = col_sum(o[t]) + b3;
            if (quad == 0 && row < nnz) out[row] = v;
        }
    }
}

// Backward activations, one 32-row chunk c per pass: row c*32 + 16h + j
// (tile h, column j) takes chunk slot 2j + h, so a lane holds two adjacent
// slots of each of its units as one 32-bit word.  Writes, as BF16 in
// unit-major order ([c][stored unit][32 slots], units in unit_store order),
// X^T (16x32), dz2^T and dz1^T (128x32 each); the wgrad kernel sums over
// slots, so their order is free.  Every BF16 weight fragment read from LDS feeds both tiles.
// Bias and w3 gradients are accumulated per lane in FP32 and summed per
// block in a fixed order.
// Two waves per SIMD: the register budget is capped at 256 (no scratch
// results), which with 77 KB of LDS per block gives two blocks per CU.
__global__ void __launch_bounds__(64 * kMlpWaves) __attribute__((amdgpu_waves_per_eu(2, 2))) k_mlp_backward(MlpArgs a, const float* __restrict__ feat,
                                                                 const float* __restrict__ go, int64_t nnz,
                                                                 uint32_t* __restrict__ xc, uint32_t* __restrict__ dz2c,
                                                                 uint32_t* __restrict__ dz1c, float* __restrict__ bias_part) {
    __shared__ uint2 w1s[kFragW1 / 4];
    __shared__ uint4 w2s[kFragW2 / 8];
    __shared__ uint4 w2t[kFragW2 / 8];
    __shared__ float4 b1s[kH / 4], b2s[kH / 4], w3s[kH / 4];
    __shared__ float red[kMlpWaves * kBiasGrads];
    for (int k = threadIdx.x; k < kH; k += blockDim.x) {
        reinterpret_cast<float*>(b1s)[k] = a.params[kOffB1 + k];
        reinterpret_cast<float*>(b2s)[k] = a.params[kOffB2 + k];
        reinterpret_cast<float*>(w3s)[k] = a.params[kOffW3 + k];
    }
    {
        const uint4* g2t = reinterpret_cast<const uint4*>(a.frag + kFragW1 + kFragW2);
        for (int k = threadIdx.x; k < kFragW2 / 8; k += blockDim.x) w2t[k] = g2t[k];
    }
    stage_frags(a.frag, kFragW1, w1s, w2s);
    const int lane = threadIdx.x & 63;
    const int quad = lane >> 4;
    const int col = lane & 15;
    const int64_t chunks = (nnz + 31) / 32;
    const int64_t wave = int64_t(blockIdx.x) * kMlpWaves + (threadIdx.x >> 6);
    const int64_t nwaves = int64_t(gridDim.x) * kMlpWaves;
    float gb1[8][4], gb2[8][4], gw3[8][4];
#pragma unroll
    for (int ob = 0; ob < 8; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) gb1[ob][r] = gb2[ob][r] = gw3[ob][r] = 0.f;
    float gb3 = 0.f;
    for (int64_t c = wave; c < chunks; c += nwaves) {
        const int64_t row0 = c * 32 + col, row1 = row0 + 16;
        const float4 xr0 = load_x_raw(feat, row0, nnz, quad);
        const float4 xr1 = load_x_raw(feat, row1, nnz, quad);
        const float g0 = row0 < nnz ? go[row0] : 0.f;
        const float g1 = row1 < nnz ? go[row1] : 0.f;
        if (quad == 0) {
            gb3 += g0;
            gb3 += g1;
        }
        // uint32 word of (stored unit s, slot pair col) in chunk c: ub + 16 s;
        // unit ob*16 + quad*4 + r is stored as ob*16 + r*4 + quad
        const int64_t ub = c * (kH * 16) + quad * 16 + col;
        const int64_t xb = c * (kF * 16) + quad * 16 + col;
#pragma unroll
        for (int r = 0; r < 4; ++r) xc[xb + r * 64] = pack2(comp(xr0, r), comp(xr1, r));
        const v4s x0 = x_bf16(xr0), x1 = x_bf16(xr1);
        // layer 1: relu(z1) packed as layer 2's B operand (not stored; the
        // wgrad kernel recomputes it); the packs also give relu'(z1) for dz1 (unit ib*16 + quad*4 + r sits in
        // slot (ib & 1) * 4 + r of pack ib / 2)
        uint4 h1b0[4], h1b1[4];
#pragma unroll
        for (int kb2 = 0; kb2 < 4; ++kb2) {
            v4f z0[2], z1[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ob = 2 * kb2 + h;
                const v4s w = as_v4s(w1s[ob * 64 + lane]);
                const float4 b = b1s[ob * 4 + quad];
                const v4f bias = {b.x, b.y, b.z, b.w};
                const v4f c0 = mfma16(w, x0, bias);
                const v4f c1 = mfma16(w, x1, bias);
                z0[h] = c0;
                z1[h] = c1;
            }
            h1b0[kb2] = pack8(z0[0], z0[1], true);
            h1b1[kb2] = pack8(z1[0], z1[1], true);
        }
        // (keeps the packs formed here rather than next to their uses, which
        // would hold the FP32 pre-activations live through layer 2)
        __builtin_amdgcn_sched_barrier(0);
        // layer 2 and dz2 = relu'(z2) * g * w3
        uint4 d2b0[4], d2b1[4];
        v4f dp0 = {0.f, 0.f, 0.f, 0.f}, dp1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ob = 0; ob < 8; ++ob) {
            const float4 b = b2s[ob * 4 + quad];
            v4f c0 = {b.x, b.y, b.z, b.w}, c1 = c0;
#pragma unroll
            for (int kb2 = 0; kb2 < 4; ++kb2) {
                const v8bf w = as_v8bf(w2s[(ob * 4 + kb2) * 64 + lane]);
                c0 = mfma32(w, as_v8bf(h1b0[kb2]), c0);
                c1 = mfma32(w, as_v8bf(h1b1[kb2]), c1);
            }
            const float4 w3 = w3s[ob * 4 + quad];
            v4f d0, d1;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                gw3[ob][r] = fmaf(g0, relu(c0[r]), gw3[ob][r]);
                gw3[ob][r] = fmaf(g1, relu(c1[r]), gw3[ob][r]);
                d0[r] = c0[r] > 0.f ? g0 * comp(w3, r) : 0.f;
                d1[r] = c1[r] > 0.f ? g1 * comp(w3, r) : 0.f;
                gb2[ob][r] += d0[r];
                gb2[ob][r] += d1[r];
                dz2c[ub + (ob * 16 + r * 4) * 16] = pack2(d0[r], d1[r]);
            }
            if (ob & 1) {
                d2b0[ob >> 1] = pack8(dp0, d0, false);
                d2b1[ob >> 1] = pack8(dp1, d1, false);
            } else {
                dp0 = d0;
                dp1 = d1;
            }
        }
        // dz1 = relu'(z1) * (W2 dz2)
#pragma unroll
        for (int ib = 0; ib < 8; ++ib) {
            v4f a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb2 = 0; kb2 < 4; ++kb2) {
                const v8bf w = as_v8bf(w2t[(ib * 4 + kb2) * 64 + lane]);
                a0 = mfma32(w, as_v8bf(d2b0[kb2]), a0);
                a1 = mfma32(w, as_v8bf(d2b1[kb2]), a1);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int slot = (ib & 1) * 4 + r;
                const float d0 = slot_nonzero(h1b0[ib >> 1], slot) ? a0[r] : 0.f;
                const float d1 = slot_nonzero(h1b1[ib >> 1], slot) ? a1[r] : 0.f;
                gb1[ib][r] += d0;
                gb1[ib][r] += d1;
                dz1c[ub + (ib * 16 + r * 4) * 16] = pack2(d0, d1);
            }
        }
    }
    // per-block partials b1 | b2 | w3 | b3: per wave over its 16 columns,
    // then the block's waves in order
    float* rw = red + (threadIdx.x >> 6) * kBiasGrads;
#pragma unroll
    for (int ob = 0; ob < 8; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float s1 = quad_sum(gb1[ob][r]);
            const float s2 = quad_sum(gb2[ob][r]);
            const float s3 = quad_sum(gw3[ob][r]);
            if (col == 0) {
                const int u = ob * 16 + quad * 4 + r;
                rw[u] = s1;
                rw[kH + u] = s2;
                rw[2 * kH + u] = s3;
            }
        }
    const float s = quad_sum(gb3);   // quad 0 lanes hold the column sums
    if (lane == 0) rw[3 * kH] = s;
    __syncthreads();
    for (int k = threadIdx.x; k < kBiasGrads; k += blockDim.x) {
        float t = 0.f;
#pragma unroll
        for (int w = 0; w < kMlpWaves; ++w) t += red[w * kBiasGrads + k];
        bias_part[int64_t(blockIdx.x) * kBiasGrads + k] = t;
    }
}

// Weight gradients by split-K MFMA over chunks: dW2 = H1^T dZ2, dW1 = X^T dZ1.
// Wave w of a block owns dW2 rows [32w, 32w+32) and dW1 columns [32w, 32w+32)
// (stored-unit order).  Operand element (stored unit, k = quad*8..+7) of a
// chunk sits at uint4 index unit*4 + quad.  H1 is not stored: each wave
// recomputes relu(X W1 + b1) for its 32 units from the FP32 features with
// X as the MFMA's A operand, which leaves lane (unit col, quad) holding rows
// 4quad + t of both tiles h, i.e. chunk slots quad*8 + 2t + h in order: the
// dW2 A operand as it stands.  The next chunk's loads are issued before this
// chunk's MFMAs, so a wave always has one chunk of loads in flight (the
// summation order over chunks is unchanged).
struct WgradOps {
    float4 x[2];
    uint4 z1[2], z2[8], xc;
};

__device__ __forceinline__ void wgrad_load(const float* __restrict__ feat, int64_t nnz, const uint16_t* __restrict__ xc,
                                           const uint16_t* __restrict__ dz2c, const uint16_t* __restrict__ dz1c,
                                           int64_t c, int w, int col, int quad, WgradOps& o) {
    const uint4* Z2 = reinterpret_cast<const uint4*>(dz2c + c * (kH * 32));
    const uint4* Z1 = reinterpret_cast<const uint4*>(dz1c + c * (kH * 32));
    const uint4* X = reinterpret_cast<const uint4*>(xc + c * (kF * 32));
#pragma unroll
    for (int h = 0; h < 2; ++h) o.x[h] = load_x_raw(feat, c * 32 + h * 16 + col, nnz, quad);
#pragma unroll
    for (int i = 0; i < 2; ++i) o.z1[i] = Z1[((2 * w + i) * 16 + col) * 4 + quad];
#pragma unroll
    for (int j = 0; j < 8; ++j) o.z2[j] = Z2[(j * 16 + col) * 4 + quad];
    o.xc = X[col * 4 + quad];
}

__global__ void __launch_bounds__(256) k_mlp_wgrad(MlpArgs a, const float* __restrict__ feat, int64_t nnz,
                                                   const uint16_t* __restrict__ xc,
                                                   const uint16_t* __restrict__ dz2c,
                                                   const uint16_t* __restrict__ dz1c, int64_t chunks,
                                                   float* __restrict__ w_part) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int quad = lane >> 4;
    const int col = lane & 15;
    // W1 columns and b1 of this lane's stored unit in each of the wave's
    // two blocks, as the B operand (features 4quad..4quad+3) and the initial
    // accumulator of the recompute
    v4s w1b[2];
    v4f b1v[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int ob = 2 * w + i;
        w1b[i] = as_v4s(reinterpret_cast<const uint2*>(a.frag)[ob * 64 + quad * 16 + unit_store(col)]);
        const float b = a.params[kOffB1 + ob * 16 + unit_store(col)];
        b1v[i] = v4f{b, b, b, b};
    }
    v4f acc2[2][8];
    v4f acc1[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        acc1[i] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 8; ++j) acc2[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    auto accumulate = [&](const WgradOps& o) {
        const v4s x0 = x_bf16(o.x[0]), x1 = x_bf16(o.x[1]);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const v4f h0 = mfma16(x0, w1b[i], b1v[i]);
            const v4f h1 = mfma16(x1, w1b[i], b1v[i]);
            v8bf ha;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                ha[2 * t] = static_cast<__bf16>(relu(h0[t]));
                ha[2 * t + 1] = static_cast<__bf16>(relu(h1[t]));
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) acc2[i][j] = mfma32(ha, as_v8bf(o.z2[j]), acc2[i][j]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) acc1[i] = mfma32(as_v8bf(o.xc), as_v8bf(o.z1[i]), acc1[i]);
    };
    // two operand buffers used in turn; a load past the last chunk reads the
    // last chunk again (in bounds, never accumulated)
    const int64_t G = gridDim.x;
    const int64_t last = chunks - 1;
    WgradOps b0, b1;
    if (int64_t(blockIdx.x) < chunks) wgrad_load(feat, nnz, xc, dz2c, dz1c, blockIdx.x, w, col, quad, b0);
    // (the scheduling barriers keep each chunk's loads ahead of the MFMAs
    // of the chunk before; the scheduler would otherwise sink every load
    // next to its first use and wait on each in turn)
    for (int64_t c = blockIdx.x; c < chunks; c += 2 * G) {
        wgrad_load(feat, nnz, xc, dz2c, dz1c, min(c + G, last), w, col, quad, b1);
        __builtin_amdgcn_sched_barrier(0);
        accumulate(b0);
        if (c + G >= chunks) break;
        wgrad_load(feat, nnz, xc, dz2c, dz1c, min(c + 2 * G, last), w, col, quad, b0);
        __builtin_amdgcn_sched_barrier(0);
        accumulate(b1);
    }
    float* out = w_part + int64_t(blockIdx.x) * (kF * kH + kH * kH);
    // dW1[f][u] at out[f*128 + u]; dW2[i][j] at out[2048 + i*128 + j]; the
    // operands' rows and columns are stored units and features, mapped back
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int ui = unit_store((2 * w + i) * 16 + quad * 4 + r);
            out[unit_store(quad * 4 + r) * kH + unit_store((2 * w + i) * 16 + col)] = acc1[i][r];
#pragma unroll
            for (int j = 0; j < 8; ++j) out[kF * kH + ui * kH + unit_store(j * 16 + col)] = acc2[i][j][r];
        }
}

// grad[p] = sum over partial blocks in fixed order, into the flat layout.
__global__ void k_grad_reduce(const float* __restrict__ w_part, int32_t nwp, const float* __restrict__ bias_part,
                              int32_t nbp, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float s = 0.f;
    if (p < kOffB1) {
        for (int b = 0; b < nwp; ++b) s += w_part[int64_t(b) * (kF * kH + kH * kH) + p];
    } else if (p >= kOffW2 && p < kOffB2) {
        for (int b = 0; b < nwp; ++b) s += w_part[int64_t(b) * (kF * kH + kH * kH) + kF * kH + (p - kOffW2)];
    } else {
        int off;
        if (p < kOffW2) off = p - kOffB1;
        else if (p < kOffW3) off = kH + (p - kOffB2);
        else if (p < kOffB3) off = 2 * kH + (p - kOffW3);
        else off = 3 * kH;
        for (int b = 0; b < nbp; ++b) s += bias_part[int64_t(b) * kBiasGrads + off];
    }
    grad[p] = s;
}

// BF16 fragments from FP32 masters.
__global__ void k_pack(const float* __restrict__ params, uint16_t* __restrict__ frag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kFrag) return;
    float v;
    if (t < kFragW1) {
        const int ob = t / 256, lane = (t / 4) % 64, r = t % 4;
        v = params[kOffW1 + ((lane >> 4) * 4 + r) * kH + ob * 16 + (lane & 15)];
    } else {
        const int u = t - kFragW1;
        const int which = u / kFragW2;
        const int e = u % kFragW2;
        const int blk = e / 512, lane = (e / 8) % 64, s = e % 8;
        const int ob = blk / 4, kb2 = blk % 4;
        const int k = unit_of_slot(kb2, lane >> 4, s);
        const int m = ob * 16 + (lane & 15);
        // forward: W2[k][m] (rows of W2^T = output unit m); backward: W2[m][k]
        v = which == 0 ? params[kOffW2 + k * kH + m] : params[kOffW2 + m * kH + k];
    }
    frag[t] = f2bf(v);
}

__global__ void k_adam(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, float lr, float beta1, float beta2, float eps, float bc1, float bc2,
                       float clip_scale) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    const float g = grad[p] * clip_scale;
    const float mm = beta1 * m[p] + (1.f - beta1) * g;
    const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

constexpr int kGradBlocks = 512;   // split-K blocks of the weight-gradient kernel (two per CU)
constexpr int kBackBlocks = 512;   // backward activation grid cap: two 4-wave blocks per CU on 256 CUs

unsigned grid_of(int64_t items, int threads, int64_t cap = 65536) {
    int64_t g = (items + threads - 1) / threads;
    return unsigned(std::max<int64_t>(1, std::min<int64_t>(g, cap)));
}

}  // namespace
}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_fsai_features(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t n,
                       const double* xyz, const int64_t* lptr, const int32_t* lidx, double* rsqrt_diag,
                       double* row_stats, float* features, void* stream) {
    SKFA_REQUIRE(indptr && indices && vals && lptr && lidx && rsqrt_diag && row_stats && features, "null argument");
    if (n == 0) return kOk;
    hipStream_t s = static_cast<hipStream_t>(stream);
    k_rsqrt_diag<<<grid_of(n, 256), 256, 0, s>>>(indptr, indices, vals, lptr, n, rsqrt_diag);
    k_row_stats<<<grid_of(n * 16, 256), 256, 0, s>>>(indptr, indices, vals, rsqrt_diag, xyz, n, row_stats);
    int64_t nnz_l = 0;
    SKFA_HIP(hipMemcpyAsync(&nnz_l, lptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    if (nnz_l)
        k_features<<<grid_of(nnz_l, 256), 256, 0, s>>>(indptr, indices, vals, rsqrt_diag, row_stats, xyz, lptr, lidx,
                                                      n, nnz_l, features);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_fsai_factor(const float* outputs, const int64_t* lptr, const int32_t* lidx, const double* rsqrt_diag,
                     int64_t n, double* gtilde, float* g, double* logdiag, void* stream) {
    SKFA_REQUIRE(outputs && lptr && lidx && rsqrt_diag, "null argument");
    if (n == 0) return kOk;
    k_gtilde<<<grid_of(n * 16, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(outputs, lptr, lidx, rsqrt_diag, n,
                                                                            gtilde, g, logdiag);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_fsai_kaporin(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t n,
                      const int64_t* sys_ptr, int32_t nsys, const int64_t* lptr, const int32_t* lidx,
                      const double* rsqrt_diag, const float* outputs, double grad_scale, double* work,
                      int32_t* row_sys, double* loss, float* grad_outputs, void* stream) {
    SKFA_REQUIRE(indptr && sys_ptr && lptr && lidx && rsqrt_diag && outputs && work && row_sys && loss,
                 "null argument");
    SKFA_REQUIRE(nsys > 0 && nsys <= 65535, "nsys must be in [1, 65535]");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t nnz_l = 0;
    SKFA_HIP(hipMemcpyAsync(&nnz_l, lptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    // work layout (doubles): gt[nnz_l] | P[nnz_l] | prod[nnz_l] | logdiag[n] | tr[nsys] | logsum[nsys] | seg[nsys+1]
    double* gt = work;
    double* P = gt + nnz_l;
    double* prod = P + nnz_l;
    double* logdiag = prod + nnz_l;
    double* tr = logdiag + n;
    double* logsum = tr + nsys;
    int64_t* seg = reinterpret_cast<int64_t*>(logsum + nsys);
    const unsigned g = grid_of(n, 256);
    k_gtilde<<<grid_of(n * 16, 256), 256, 0, s>>>(outputs, lptr, lidx, rsqrt_diag, n, gt, nullptr, logdiag);
    k_kaporin_p<<<g, 256, 0, s>>>(indptr, indices, vals, rsqrt_diag, lptr, lidx, gt, n, P, prod);
    k_seg_of<<<(nsys + 256) / 256, 256, 0, s>>>(sys_ptr, lptr, nsys, seg);
    k_segsum<<<nsys, kRed, 0, s>>>(prod, seg, tr);
    k_segsum<<<nsys, kRed, 0, s>>>(logdiag, sys_ptr, logsum);
    k_kaporin_loss<<<(nsys + 255) / 256, 256, 0, s>>>(tr, logsum, sys_ptr, nsys, loss);
    if (grad_outputs) {
        k_row_sys<<<dim3(32, unsigned(nsys)), 256, 0, s>>>(sys_ptr, nsys, row_sys);
        k_kaporin_grad<<<g, 256, 0, s>>>(lptr, lidx, gt, P, tr, row_sys, sys_ptr, n, grad_scale, grad_outputs);
    }
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_mlp_pack(const float* params, uint16_t* frag, void* stream) {
    SKFA_REQUIRE(params && frag, "null argument");
    k_pack<<<(kFrag + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(params, frag);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_mlp_forward(const float* params, const uint16_t* frag, const float* features, int64_t nnz, float* out,
                     void* stream) {
    SKFA_REQUIRE(params && frag && features && out, "null argument");
    if (nnz == 0) return kOk;
    const int64_t groups = (nnz + 16 * kFwdTiles - 1) / (16 * kFwdTiles);
    const unsigned blocks = unsigned(std::min<int64_t>((groups + kMlpWaves - 1) / kMlpWaves, 2048));
    k_mlp_forward<<<blocks, 64 * kMlpWaves, 0, static_cast<hipStream_t>(stream)>>>(MlpArgs{params, frag}, features,
                                                                                    nnz, out);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

size_t skfa_mlp_workspace_bytes(int64_t nnz) {
    const int64_t chunks = (nnz + 31) / 32;
    return size_t(chunks) * 32 * (kF + 2 * kH) * sizeof(uint16_t) +
           size_t(kGradBlocks) * (kF * kH + kH * kH) * sizeof(float) +
           size_t(kBackBlocks) * kBiasGrads * sizeof(float) + 256;
}

int skfa_mlp_backward(const float* params, const uint16_t* frag, const float* features, const float* grad_out,
                      int64_t nnz, void* workspace, size_t workspace_bytes, float* grad, void* stream) {
    SKFA_REQUIRE(params && frag && features && grad_out && workspace && grad, "null argument");
    SKFA_REQUIRE(workspace_bytes >= skfa_mlp_workspace_bytes(nnz), "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t chunks = (nnz + 31) / 32;
    uint16_t* xc = static_cast<uint16_t*>(workspace);
    uint16_t* dz2c = xc + chunks * 32 * kF;
    uint16_t* dz1c = dz2c + chunks * 32 * kH;
    uintptr_t pw = reinterpret_cast<uintptr_t>(dz1c + chunks * 32 * kH);
    pw = (pw + 255) & ~uintptr_t(255);
    float* w_part = reinterpret_cast<float*>(pw);
    float* bias_part = w_part + size_t(kGradBlocks) * (kF * kH + kH * kH);
    // one chunk per wave up to the cap; the grid depends on nnz only, so the
    // bias partials (one per block) and their sum order are reproducible
    const int bblocks = int(std::max<int64_t>(1, std::min<int64_t>((chunks + kMlpWaves - 1) / kMlpWaves, kBackBlocks)));
    k_mlp_backward<<<bblocks, 64 * kMlpWaves, 0, s>>>(
        MlpArgs{params, frag}, features, grad_out, nnz, reinterpret_cast<uint32_t*>(xc),
        reinterpret_cast<uint32_t*>(dz2c), reinterpret_cast<uint32_t*>(dz1c), bias_part);
    SKFA_LAUNCH_CHECK();
    // at least eight chunks per block, so small batches write few partials
    const int wblocks = int(std::max<int64_t>(1, std::min<int64_t>((chunks + 7) / 8, kGradBlocks)));
    k_mlp_wgrad<<<wblocks, 256, 0, s>>>(MlpArgs{params, frag}, features, nnz, xc, dz2c, dz1c, chunks, w_part);
    SKFA_LAUNCH_CHECK();
    k_grad_reduce<<<(kParams + 255) / 256, 256, 0, s>>>(w_part, wblocks, bias_part, bblocks, grad);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                  float eps, int32_t step, float clip_scale, uint16_t* frag, void* stream) {
    SKFA_REQUIRE(params && grad && m && v && frag && step >= 1, "invalid argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const float bc1 = 1.f - powf(beta1, float(step));
    const float bc2 = 1.f - powf(beta2, float(step));
    k_adam<<<(kParams + 255) / 256, 256, 0, s>>>(params, grad, m, v, lr, beta1, beta2, eps, bc1, bc2, clip_scale);
    k_pack<<<(kFrag + 255) / 256, 256, 0, s>>>(params, frag);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

}  // extern "C"
