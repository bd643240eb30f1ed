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

__global__ void k_rsqrt_diag(const int64_t* __restrict__ ip, const int32_t* __restrict__ ix,
                             const double* __restrict__ av, int64_t n, double* __restrict__ rs) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        double d = 0.;
        for (int64_t k = ip[i]; k < ip[i + 1]; ++k)
            if (ix[k] == i) d = av[k];
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
constexpr int kFeatBlock = 256;

__global__ void __launch_bounds__(kFeatBlock) k_features(const int64_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                                         const double* __restrict__ av, const double* __restrict__ rs,
                                                         const double* __restrict__ stats, const double* __restrict__ xyz,
                                                         const int64_t* __restrict__ lptr, const int32_t* __restrict__ lidx,
                                                         int64_t n, int64_t nnz_l, float* __restrict__ feat) {
    __shared__ int64_t win[kFeatBlock + 1];
    __shared__ int64_t first_row;
    for (int64_t q0 = int64_t(blockIdx.x) * kFeatBlock; q0 < nnz_l; q0 += int64_t(gridDim.x) * kFeatBlock) {
        if (threadIdx.x == 0) first_row = row_of(lptr, n, q0);
        __syncthreads();
        const int64_t i0 = first_row;
        // win[k] = lptr[i0 + k]; rows past n read as nnz_l
        for (int k = threadIdx.x; k <= kFeatBlock; k += blockDim.x) win[k] = i0 + k <= n ? lptr[i0 + k] : nnz_l;
        __syncthreads();
        const int64_t q = q0 + threadIdx.x;
        int lo = 0, hi = kFeatBlock;   // largest k with win[k] <= q
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (win[mid] <= q) lo = mid;
            else hi = mid;
        }
        __syncthreads();
        if (q >= nnz_l) continue;
        const int64_t i = i0 + lo;
        const int32_t j = lidx[q];
        const int64_t bi = ip[i], ei = ip[i + 1];
        const int64_t bj = ip[j], ej = ip[j + 1];
        const double rsi = rs[i], rsj = rs[j];
        double aij = 0., t_mid = 0., t_lo = 0., t_all = 0.;
        // sorted merge of rows i and j
        int64_t a = bi, b = bj;
        while (a < ei && b < ej) {
            const int32_t ca = ix[a], cb = ix[b];
            if (ca == cb) {
                const double rk = rs[ca];
                const double p = (av[a] * rsi * rk) * (av[b] * rsj * rk);
                t_all += p;
                if (ca < j) t_lo += p;
                else if (ca > j && ca < i) t_mid += p;
                if (ca == j) aij = av[a] * rsi * rsj;
                ++a;
                ++b;
            } else if (ca < cb) {
                ++a;
            } else {
                ++b;
            }
        }
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
}

// ------------------------------------------------------------------ Kaporin objective
// Gt from network outputs; G = Gt D^{-1/2} as float; per-nonzero diag flag.
__global__ void k_gtilde(const float* __restrict__ o, const int64_t* __restrict__ lptr,
                         const int32_t* __restrict__ lidx, const double* __restrict__ rs, int64_t n,
                         double* __restrict__ gt, float* __restrict__ g, double* __restrict__ logdiag) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        for (int64_t q = lptr[i]; q < lptr[i + 1]; ++q) {
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
__device__ __forceinline__ v4s load_x(const float* __restrict__ feat, int64_t row, int64_t nnz, int quad) {
    return x_bf16(load_x_raw(feat, row, nnz, quad));
}

// Forward of one tile.  Outputs per lane (units ob*16 + quad*4 + r, column
// lane%16): z1[8] and z2[8] pre-activations (FP32), h1b[4] BF16 packs of
// relu(z1) in the 32-deep B-operand order.
struct TileFwd {
    v4f z1[8];
    v4f z2[8];
    uint4 h1b[4];
};

__device__ __forceinline__ uint4 pack8(const v4f& a, const v4f& b, bool relu) {
    const float f[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
    v8bf h;
#pragma unroll
    for (int s = 0; s < 8; ++s) h[s] = static_cast<__bf16>(relu ? fmaxf(f[s], 0.f) : f[s]);
    return __builtin_bit_cast(uint4, h);
}

__device__ __forceinline__ void tile_forward(const uint2* __restrict__ w1f, const uint4* __restrict__ w2f,
                                             const float* __restrict__ params, v4s x, int lane, TileFwd& t) {
    const int quad = lane >> 4;
#pragma unroll
    for (int ob = 0; ob < 8; ++ob) {
        v4f c = {0.f, 0.f, 0.f, 0.f};
        c = mfma16(as_v4s(w1f[ob * 64 + lane]), x, c);
#pragma unroll
        for (int r = 0; r < 4; ++r) c[r] += params[kOffB1 + ob * 16 + quad * 4 + r];
        t.z1[ob] = c;
    }
#pragma unroll
    for (int kb2 = 0; kb2 < 4; ++kb2) t.h1b[kb2] = pack8(t.z1[2 * kb2], t.z1[2 * kb2 + 1], true);
#pragma unroll
    for (int ob = 0; ob < 8; ++ob) {
        v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb2 = 0; kb2 < 4; ++kb2) c = mfma32(as_v8bf(w2f[(ob * 4 + kb2) * 64 + lane]), as_v8bf(t.h1b[kb2]), c);
#pragma unroll
        for (int r = 0; r < 4; ++r) c[r] += params[kOffB2 + ob * 16 + quad * 4 + r];
        t.z2[ob] = c;
    }
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

// Inference form of tile_forward: the output unit's dot product is taken
// right after each 16-unit block of layer 2 (same summation order as
// summing the full z2 afterwards), so only one accumulator tile is live.
// b1, b2 and w3 come from LDS (float4 per lane: units ob*16 + quad*4 + 0..3).
__device__ __forceinline__ float tile_forward_out(const uint2* __restrict__ w1f, const uint4* __restrict__ w2f,
                                                  const float4* __restrict__ b1s, const float4* __restrict__ b2s,
                                                  const float4* __restrict__ w3s, v4s x, int lane) {
    const int quad = lane >> 4;
    uint4 h1b[4];
#pragma unroll
    for (int kb2 = 0; kb2 < 4; ++kb2) {
        v4f z[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ob = 2 * kb2 + h;
            v4f c = {0.f, 0.f, 0.f, 0.f};
            c = mfma16(as_v4s(w1f[ob * 64 + lane]), x, c);
            const float4 b = b1s[ob * 4 + quad];
            c[0] += b.x;
            c[1] += b.y;
            c[2] += b.z;
            c[3] += b.w;
            z[h] = c;
        }
        h1b[kb2] = pack8(z[0], z[1], true);
    }
    float o = 0.f;
#pragma unroll
    for (int ob = 0; ob < 8; ++ob) {
        v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb2 = 0; kb2 < 4; ++kb2) c = mfma32(as_v8bf(w2f[(ob * 4 + kb2) * 64 + lane]), as_v8bf(h1b[kb2]), c);
        const float4 b = b2s[ob * 4 + quad];
        const float4 w = w3s[ob * 4 + quad];
        c[0] += b.x;
        c[1] += b.y;
        c[2] += b.z;
        c[3] += b.w;
        o = fmaf(w.x, fmaxf(c[0], 0.f), o);
        o = fmaf(w.y, fmaxf(c[1], 0.f), o);
        o = fmaf(w.z, fmaxf(c[2], 0.f), o);
        o = fmaf(w.w, fmaxf(c[3], 0.f), o);
        // keep one block's fragments and accumulator live at a time
        __builtin_amdgcn_sched_barrier(0);
    }
    return o;
}

__global__ void __launch_bounds__(64 * kMlpWaves) k_mlp_forward(MlpArgs a, const float* __restrict__ feat,
                                                                int64_t nnz, float* __restrict__ out) {
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
    const int64_t tiles = (nnz + 15) / 16;
    const int64_t wave = int64_t(blockIdx.x) * kMlpWaves + (threadIdx.x >> 6);
    const int64_t nwaves = int64_t(gridDim.x) * kMlpWaves;
    const float b3 = a.params[kOffB3];
    // the next tile's features are in flight while this tile is computed
    float4 xn = load_x_raw(feat, wave * 16 + (lane & 15), nnz, quad);
    for (int64_t tile = wave; tile < tiles; tile += nwaves) {
        const int64_t row = tile * 16 + (lane & 15);
        const float4 x = xn;
        xn = load_x_raw(feat, (tile + nwaves) * 16 + (lane & 15), nnz, quad);
        float o = tile_forward_out(w1s, w2s, b1s, b2s, w3s, x_bf16(x), lane);
        o = col_sum(o) + b3;
        if (quad == 0 && row < nnz) out[row] = o;
    }
}

// Backward activations.  For each 32-row chunk c (two tiles) writes, as BF16
// in unit-major order, X^T (16x32), relu(z1)^T, dz2^T and dz1^T (128x32
// each); bias and w3 gradients are accumulated per wave.
__global__ void __launch_bounds__(64 * kMlpWaves) k_mlp_backward(MlpArgs a, const float* __restrict__ feat,
                                                                 const float* __restrict__ go, int64_t nnz,
                                                                 uint16_t* __restrict__ xc, uint16_t* __restrict__ h1c,
                                                                 uint16_t* __restrict__ dz2c, uint16_t* __restrict__ dz1c,
                                                                 float* __restrict__ bias_part) {
    __shared__ uint2 w1s[kFragW1 / 4];
    __shared__ uint4 w2s[kFragW2 / 8];
    stage_frags(a.frag, kFragW1, w1s, w2s);
    const uint4* w2b = reinterpret_cast<const uint4*>(a.frag + kFragW1 + kFragW2);
    const int lane = threadIdx.x & 63;
    const int quad = lane >> 4;
    const int col = lane & 15;
    const int64_t chunks = (nnz + 31) / 32;
    const int64_t wave = int64_t(blockIdx.x) * kMlpWaves + (threadIdx.x >> 6);
    const int64_t nwaves = int64_t(gridDim.x) * kMlpWaves;
    float w3[8][4];
#pragma unroll
    for (int ob = 0; ob < 8; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) w3[ob][r] = a.params[kOffW3 + ob * 16 + quad * 4 + r];
    float gb1[8][4], gb2[8][4], gw3[8][4];
#pragma unroll
    for (int ob = 0; ob < 8; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) gb1[ob][r] = gb2[ob][r] = gw3[ob][r] = 0.f;
    float gb3 = 0.f;
    for (int64_t c = wave; c < chunks; c += nwaves) {
#pragma unroll 1
        for (int half = 0; half < 2; ++half) {
            const int64_t row = c * 32 + half * 16 + col;
            const v4s x = load_x(feat, row, nnz, quad);
            TileFwd t;
            tile_forward(w1s, w2s, a.params, x, lane, t);
            const float g = (row < nnz) ? go[row] : 0.f;
            if (quad == 0) gb3 += g;
            v4f dz2[8];
#pragma unroll
            for (int ob = 0; ob < 8; ++ob)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float h2 = fmaxf(t.z2[ob][r], 0.f);
                    gw3[ob][r] = fmaf(g, h2, gw3[ob][r]);
                    const float d = t.z2[ob][r] > 0.f ? g * w3[ob][r] : 0.f;
                    dz2[ob][r] = d;
                    gb2[ob][r] += d;
                }
            uint4 dz2b[4];
#pragma unroll
            for (int kb2 = 0; kb2 < 4; ++kb2) dz2b[kb2] = pack8(dz2[2 * kb2], dz2[2 * kb2 + 1], false);
            v4f dz1[8];
#pragma unroll
            for (int ib = 0; ib < 8; ++ib) {
                v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb2 = 0; kb2 < 4; ++kb2) acc = mfma32(as_v8bf(w2b[(ib * 4 + kb2) * 64 + lane]), as_v8bf(dz2b[kb2]), acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float d = t.z1[ib][r] > 0.f ? acc[r] : 0.f;
                    dz1[ib][r] = d;
                    gb1[ib][r] += d;
                }
            }
            // chunk-transposed BF16 stores: [c][unit][32]
            const int64_t base = c * (kH * 32) + half * 16 + col;
#pragma unroll
            for (int ob = 0; ob < 8; ++ob)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int u = ob * 16 + quad * 4 + r;
                    h1c[base + u * 32] = f2bf(fmaxf(t.z1[ob][r], 0.f));
                    dz2c[base + u * 32] = f2bf(dz2[ob][r]);
                    dz1c[base + u * 32] = f2bf(dz1[ob][r]);
                }
            const int64_t xbase = c * (kF * 32) + half * 16 + col;
            const uint2 xu = __builtin_bit_cast(uint2, x);
            xc[xbase + (quad * 4 + 0) * 32] = uint16_t(xu.x & 0xffffu);
            xc[xbase + (quad * 4 + 1) * 32] = uint16_t(xu.x >> 16);
            xc[xbase + (quad * 4 + 2) * 32] = uint16_t(xu.y & 0xffffu);
            xc[xbase + (quad * 4 + 3) * 32] = uint16_t(xu.y >> 16);
        }
    }
    // per-wave partials: b1 | b2 | w3 | b3
    float* out = bias_part + wave * kBiasGrads;
#pragma unroll
    for (int ob = 0; ob < 8; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float s1 = quad_sum(gb1[ob][r]);
            const float s2 = quad_sum(gb2[ob][r]);
            const float s3 = quad_sum(gw3[ob][r]);
            if (col == 0) {
                const int u = ob * 16 + quad * 4 + r;
                out[u] = s1;
                out[kH + u] = s2;
                out[2 * kH + u] = s3;
            }
        }
    const float s = quad_sum(gb3);   // quad 0 lanes hold the column sums
    if (lane == 0) out[3 * kH] = s;
}

// Weight gradients by split-K MFMA over chunks: dW2 = H1^T dZ2, dW1 = X^T dZ1.
// Wave w of a block owns dW2 rows [32w, 32w+32) and dW1 columns [32w, 32w+32).
__global__ void __launch_bounds__(256) k_mlp_wgrad(const uint16_t* __restrict__ xc, const uint16_t* __restrict__ h1c,
                                                   const uint16_t* __restrict__ dz2c,
                                                   const uint16_t* __restrict__ dz1c, int64_t chunks,
                                                   float* __restrict__ w_part) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int quad = lane >> 4;
    const int col = lane & 15;
    v4f acc2[2][8];
    v4f acc1[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        acc1[i] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 8; ++j) acc2[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    for (int64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const uint4* H = reinterpret_cast<const uint4*>(h1c + c * (kH * 32));
        const uint4* Z2 = reinterpret_cast<const uint4*>(dz2c + c * (kH * 32));
        const uint4* Z1 = reinterpret_cast<const uint4*>(dz1c + c * (kH * 32));
        const uint4* Xc = reinterpret_cast<const uint4*>(xc + c * (kF * 32));
        // operand element (unit, k=quad*8..+7) sits at uint4 index unit*4 + quad
        v8bf a2[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a2[i] = as_v8bf(H[((2 * w + i) * 16 + col) * 4 + quad]);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const v8bf b = as_v8bf(Z2[(j * 16 + col) * 4 + quad]);
#pragma unroll
            for (int i = 0; i < 2; ++i) acc2[i][j] = mfma32(a2[i], b, acc2[i][j]);
        }
        const v8bf ax = as_v8bf(Xc[col * 4 + quad]);
#pragma unroll
        for (int i = 0; i < 2; ++i) acc1[i] = mfma32(ax, as_v8bf(Z1[((2 * w + i) * 16 + col) * 4 + quad]), acc1[i]);
    }
    float* out = w_part + int64_t(blockIdx.x) * (kF * kH + kH * kH);
    // dW1[f][u] at out[f*128 + u]; dW2[i][j] at out[2048 + i*128 + j]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            out[(quad * 4 + r) * kH + (2 * w + i) * 16 + col] = acc1[i][r];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                out[kF * kH + ((2 * w + i) * 16 + quad * 4 + r) * kH + j * 16 + col] = acc2[i][j][r];
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

constexpr int kGradBlocks = 256;

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
    const unsigned g = grid_of(n, 256);
    k_rsqrt_diag<<<g, 256, 0, s>>>(indptr, indices, vals, n, rsqrt_diag);
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
    k_gtilde<<<grid_of(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(outputs, lptr, lidx, rsqrt_diag, n,
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
    k_gtilde<<<g, 256, 0, s>>>(outputs, lptr, lidx, rsqrt_diag, n, gt, nullptr, logdiag);
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
    const int64_t tiles = (nnz + 15) / 16;
    const unsigned blocks = unsigned(std::min<int64_t>((tiles + kMlpWaves - 1) / kMlpWaves, 2048));
    k_mlp_forward<<<blocks, 64 * kMlpWaves, 0, static_cast<hipStream_t>(stream)>>>(MlpArgs{params, frag}, features,
                                                                                    nnz, out);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

size_t skfa_mlp_workspace_bytes(int64_t nnz) {
    const int64_t chunks = (nnz + 31) / 32;
    return size_t(chunks) * 32 * (kF + 3 * kH) * sizeof(uint16_t) +
           size_t(kGradBlocks) * (kF * kH + kH * kH) * sizeof(float) +
           size_t(kGradBlocks) * kMlpWaves * kBiasGrads * sizeof(float) + 256;
}

int skfa_mlp_backward(const float* params, const uint16_t* frag, const float* features, const float* grad_out,
                      int64_t nnz, void* workspace, size_t workspace_bytes, float* grad, void* stream) {
    SKFA_REQUIRE(params && frag && features && grad_out && workspace && grad, "null argument");
    SKFA_REQUIRE(workspace_bytes >= skfa_mlp_workspace_bytes(nnz), "workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t chunks = (nnz + 31) / 32;
    uint16_t* xc = static_cast<uint16_t*>(workspace);
    uint16_t* h1c = xc + chunks * 32 * kF;
    uint16_t* dz2c = h1c + chunks * 32 * kH;
    uint16_t* dz1c = dz2c + chunks * 32 * kH;
    uintptr_t pw = reinterpret_cast<uintptr_t>(dz1c + chunks * 32 * kH);
    pw = (pw + 255) & ~uintptr_t(255);
    float* w_part = reinterpret_cast<float*>(pw);
    float* bias_part = w_part + size_t(kGradBlocks) * (kF * kH + kH * kH);
    const int bblocks = kGradBlocks;   // backward grid: fixed, so bias partials are fixed too
    k_mlp_backward<<<bblocks, 64 * kMlpWaves, 0, s>>>(MlpArgs{params, frag}, features, grad_out, nnz, xc, h1c,
                                                      dz2c, dz1c, bias_part);
    SKFA_LAUNCH_CHECK();
    const int wblocks = int(std::max<int64_t>(1, std::min<int64_t>(chunks, kGradBlocks)));
    k_mlp_wgrad<<<wblocks, 256, 0, s>>>(xc, h1c, dz2c, dz1c, chunks, w_part);
    SKFA_LAUNCH_CHECK();
    k_grad_reduce<<<(kParams + 255) / 256, 256, 0, s>>>(w_part, wblocks, bias_part, bblocks * kMlpWaves, grad);
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
