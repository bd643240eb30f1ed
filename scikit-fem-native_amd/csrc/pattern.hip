// Device construction of the CSR pattern and of the element->CSR gather
// plan for continuous nodal spaces, plus first-seen edge numbering.
//
// Replaces, with identical integer outputs:
//  * the pattern + scatter map of BilinearFormAssembler / CrossBilinearAssembler
//    (cpp/src/assembly/bilinear_form_assembler.cpp:111-120,
//     cross_bilinear_assembler.cpp:307-320) and build_edge_csr_pattern
//    (cpp/src/assembly/edge_topology.cpp:86-154): per row, the sorted unique
//    set of column DOFs of all elements touching that row;
//  * build_oriented_edge_topology (cpp/src/assembly/edge_topology.cpp:30-84)
//    and the edge-node numbering of MeshTet2.from_mesh
//    (python/skfemntv/basis.py:644-658): edges numbered in order of first
//    occurrence over (element, local edge).
//
// The reference's greedy element colouring has no counterpart: assembly is
// a row-owner gather (see assemble.hip), so no colouring is needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "skfa_internal.h"
#include "skfa_plan.h"
#include "../../include/skfem_amd.h"

namespace skfa {

// ------------------------------------------------------------------ scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <class TIn>
__global__ void __launch_bounds__(kScanThreads) k_scan_tile(const TIn* __restrict__ in, int64_t* __restrict__ out,
                                                            int64_t n, int64_t* __restrict__ tile_sums) {
    __shared__ int64_t sh[kScanThreads];
    const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    int64_t v[kScanItems];
    int64_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        v[k] = i < n ? int64_t(in[i]) : 0;
        local += v[k];
    }
    sh[threadIdx.x] = local;
    __syncthreads();
    // Hillis-Steele inclusive scan over thread totals.
    for (int off = 1; off < kScanThreads; off <<= 1) {
        int64_t add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t run = sh[threadIdx.x] - local;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = sh[threadIdx.x];
}

__global__ void k_scan_add(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ tile_offsets) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tile_offsets[i / kScanTile];
}

__global__ void k_write_total(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ last_in_is_i64,
                              const int32_t* __restrict__ last_in_is_i32) {
    // out[n] = out[n-1] + in[n-1]
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t last = last_in_is_i64 ? last_in_is_i64[n - 1] : int64_t(last_in_is_i32[n - 1]);
        out[n] = out[n - 1] + last;
    }
}

template <class TIn>
static int scan_impl(const TIn* in, int64_t* out, int64_t n, hipStream_t s) {
    if (n == 0) {
        SKFA_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), s));
        return kOk;
    }
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    DevBuf<int64_t> sums, offs;
    if (int rc = sums.alloc(tiles)) return rc;
    if (int rc = offs.alloc(tiles + 1)) return rc;
    k_scan_tile<TIn><<<dim3(unsigned(tiles)), dim3(kScanThreads), 0, s>>>(in, out, n, sums.p);
    SKFA_LAUNCH_CHECK();
    if (tiles > 1) {
        if (int rc = scan_impl<int64_t>(sums.p, offs.p, tiles, s)) return rc;
        k_scan_add<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(out, n, offs.p);
        SKFA_LAUNCH_CHECK();
    }
    if (std::is_same<TIn, int64_t>::value)
        k_write_total<<<1, 1, 0, s>>>(out, n, reinterpret_cast<const int64_t*>(in), nullptr);
    else
        k_write_total<<<1, 1, 0, s>>>(out, n, nullptr, reinterpret_cast<const int32_t*>(in));
    SKFA_LAUNCH_CHECK();
    // the temporaries are freed on return: wait for the kernels that use them
    SKFA_HIP(hipStreamSynchronize(s));
    return kOk;
}

int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s) { return scan_impl(in, out, n, s); }
int exclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, hipStream_t s) {
    return scan_impl(in, out, n, s);
}

// ------------------------------------------------------------------ incidence
__global__ void k_inc_count(const int32_t* __restrict__ conn, int64_t total, int64_t nnodes,
                            int32_t* __restrict__ count, int32_t* __restrict__ status) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
        const int32_t v = conn[i];
        if (v < 0 || v >= nnodes) {
            atomicOr(status, 1);
            continue;
        }
        atomicAdd(&count[v], 1);
    }
}

__global__ void k_inc_fill(const int32_t* __restrict__ conn, int64_t E, int n, int64_t nnodes,
                           const int64_t* __restrict__ inc_ptr, int32_t* __restrict__ cursor,
                           int32_t* __restrict__ inc) {
    const int64_t total = E * n;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
        const int32_t v = conn[i];
        if (v < 0 || v >= nnodes) continue;
        const int64_t e = i / n;
        const int a = int(i - e * n);
        const int32_t slot = atomicAdd(&cursor[v], 1);
        inc[inc_ptr[v] + slot] = int32_t((e << 4) | a);
    }
}

// Sort each node's incidence segment (atomic fill order is arbitrary; the
// packed key (e<<4)|a orders by element then local node).
__global__ void k_inc_sort(const int64_t* __restrict__ inc_ptr, int32_t* __restrict__ inc, int64_t nnodes) {
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nnodes; r += int64_t(gridDim.x) * blockDim.x) {
        const int64_t b = inc_ptr[r], e = inc_ptr[r + 1];
        for (int64_t i = b + 1; i < e; ++i) {
            const int32_t key = inc[i];
            int64_t j = i - 1;
            while (j >= b && inc[j] > key) {
                inc[j + 1] = inc[j];
                --j;
            }
            inc[j + 1] = key;
        }
    }
}

// ------------------------------------------------------------------ rows
// One 64-lane wave per node row.  Candidates = columns (nodes of `conn`, n
// per element) of every element touching the row; for a rectangular plan
// `conn` is the column field's connectivity while the incidences come from
// the row field's. a candidate is "first" if no earlier candidate has the
// same column; its rank is the number of first candidates with a smaller
// column, which is both its CSR position within the row (sorted unique
// order, as std::sort+std::unique gives) and the element->CSR offset.
__global__ void __launch_bounds__(64) k_row_analyze(const int32_t* __restrict__ conn, int n,
                                                    const int64_t* __restrict__ inc_ptr,
                                                    const int32_t* __restrict__ inc, int64_t nnodes, int phase,
                                                    int32_t* __restrict__ row_len, const int64_t* __restrict__ indptr,
                                                    int32_t* __restrict__ indices,
                                                    const int64_t* __restrict__ row_rec,
                                                    const int32_t* __restrict__ row_stride,
                                                    uint8_t* __restrict__ records, int rec_bytes, int off_bytes,
                                                    int32_t* __restrict__ status) {
    __shared__ int32_t cand[kMaxCandidates];
    __shared__ uint8_t first[kMaxCandidates];
    const int lane = threadIdx.x;
    for (int64_t r = blockIdx.x; r < nnodes; r += gridDim.x) {
        const int64_t k0 = inc_ptr[r], k1 = inc_ptr[r + 1];
        const int64_t m64 = (k1 - k0) * n;
        if (m64 > kMaxCandidates) {
            if (lane == 0) atomicOr(status, 2);
            continue;
        }
        const int m = int(m64);
        for (int i = lane; i < m; i += 64) {
            const int k = i / n, b = i - (i / n) * n;
            const int32_t e = inc[k0 + k] >> 4;
            cand[i] = conn[int64_t(e) * n + b];
        }
        __syncthreads();
        int count = 0;
        for (int i = lane; i < m; i += 64) {
            const int32_t c = cand[i];
            bool f = true;
            for (int j = 0; j < i; ++j)
                if (cand[j] == c) { f = false; break; }
            first[i] = f;
            count += f;
        }
        __syncthreads();
        if (phase == 0) {
            // wave reduction of the first-candidate count
            for (int off = 32; off > 0; off >>= 1) count += __shfl_down(count, off, 64);
            if (lane == 0) row_len[r] = count;
        } else {
            const int64_t base = indptr[r];
            const int64_t rrec = row_rec[r];
            const int64_t rstride = row_stride ? row_stride[r] : 1;
            for (int i = lane; i < m; i += 64) {
                const int32_t c = cand[i];
                int rank = 0;
                for (int j = 0; j < m; ++j) rank += (first[j] && cand[j] < c);
                if (first[i]) indices[base + rank] = c;
                const int k = i / n, b = i - (i / n) * n;
                uint8_t* rec = records + (rrec + int64_t(k) * rstride) * rec_bytes;
                if (off_bytes == 1)
                    rec[4 + b] = uint8_t(rank);
                else
                    reinterpret_cast<uint16_t*>(rec + 4)[b] = uint16_t(rank);
                if (b == 0) *reinterpret_cast<int32_t*>(rec) = inc[k0 + k];
            }
        }
        __syncthreads();
    }
}

__global__ void k_max_diff_i64(const int64_t* __restrict__ ptr, int64_t n, int32_t* __restrict__ out) {
    int32_t mx = 0;
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += int64_t(gridDim.x) * blockDim.x)
        mx = max(mx, int32_t(ptr[r + 1] - ptr[r]));
    atomicMax(out, mx);
}

static unsigned grid_for(int64_t items, int threads, int64_t cap = 65536) {
    int64_t g = (items + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

static int plan_build(skfa_plan* P, const int32_t* conn, hipStream_t s) {
    const int64_t E = P->E, nn = P->nnodes;
    const int n = P->n;
    DevBuf<int32_t> cnt, status;
    if (int rc = cnt.alloc(nn + 1)) return rc;
    if (int rc = status.alloc(2)) return rc;
    SKFA_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (nn + 1), s));
    SKFA_HIP(hipMemsetAsync(status.p, 0, sizeof(int32_t) * 2, s));
    k_inc_count<<<grid_for(E * n, 256), 256, 0, s>>>(conn, E * n, nn, cnt.p, status.p);
    SKFA_LAUNCH_CHECK();
    if (int rc = P->inc_ptr.alloc(nn + 1)) return rc;
    if (int rc = exclusive_scan_i32_to_i64(cnt.p, P->inc_ptr.p, nn, s)) return rc;
    int32_t st = 0;
    SKFA_HIP(hipMemcpyAsync(&st, status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    if (st & 1) {
        set_error("element node index is out of bounds");
        return kInvalid;
    }
    if (int rc = P->inc.alloc(size_t(E * n))) return rc;
    SKFA_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (nn + 1), s));
    k_inc_fill<<<grid_for(E * n, 256), 256, 0, s>>>(conn, E, n, nn, P->inc_ptr.p, cnt.p, P->inc.p);
    SKFA_LAUNCH_CHECK();
    k_inc_sort<<<grid_for(nn, 256), 256, 0, s>>>(P->inc_ptr.p, P->inc.p, nn);
    SKFA_LAUNCH_CHECK();

    // row lengths -> indptr
    DevBuf<int32_t> rowlen;
    if (int rc = rowlen.alloc(nn + 1)) return rc;
    const unsigned rows_grid = unsigned(std::min<int64_t>(std::max<int64_t>(nn, 1), 1 << 20));
    k_row_analyze<<<rows_grid, 64, 0, s>>>(conn, n, P->inc_ptr.p, P->inc.p, nn, 0, rowlen.p, nullptr, nullptr,
                                           nullptr, nullptr, nullptr, 0, 1, status.p);
    SKFA_LAUNCH_CHECK();
    SKFA_HIP(hipMemcpyAsync(&st, status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    if (st & 2) {
        set_error("a node row touches more than 4096 element columns");
        return kCapacity;
    }
    if (int rc = P->indptr.alloc(nn + 1)) return rc;
    if (int rc = exclusive_scan_i32_to_i64(rowlen.p, P->indptr.p, nn, s)) return rc;
    SKFA_HIP(hipMemcpyAsync(&P->nnz_node, P->indptr.p + nn, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    // longest row / largest incidence (for the offset width and for reporting)
    SKFA_HIP(hipMemsetAsync(status.p, 0, sizeof(int32_t) * 2, s));
    k_max_diff_i64<<<grid_for(nn, 256, 1024), 256, 0, s>>>(P->indptr.p, nn, status.p);
    k_max_diff_i64<<<grid_for(nn, 256, 1024), 256, 0, s>>>(P->inc_ptr.p, nn, status.p + 1);
    SKFA_LAUNCH_CHECK();
    int32_t mx[2];
    SKFA_HIP(hipMemcpyAsync(mx, status.p, sizeof(mx), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    P->max_row_len = mx[0];
    P->max_inc = mx[1];
    if (P->nnz_node > INT32_MAX) {
        set_error("node-level pattern exceeds 2^31 nonzeros");
        return kCapacity;
    }
    if (int rc = P->indices.alloc(size_t(P->nnz_node))) return rc;

    // wave layout (host greedy partition over node rows)
    std::vector<int64_t> hptr(size_t(nn + 1)), hinc(size_t(nn + 1));
    SKFA_HIP(hipMemcpyAsync(hptr.data(), P->indptr.p, sizeof(int64_t) * (nn + 1), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipMemcpyAsync(hinc.data(), P->inc_ptr.p, sizeof(int64_t) * (nn + 1), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    const int cc = P->C * P->C;
    // lanes per node row and part: C*C (one lane per component pair), or 10
    // for 10-node 3-component fields, whose gather gives each column node
    // of an element row its own lane
    // (4 for the vector-P2 MFMA gather: a row is one MFMA column, its 4
    // lanes the 4 slices of the contraction, so a wave holds 16 rows)
    const int unit = (P->C == 3 && P->n == 10) ? (P->row_lanes == 4 ? 4 : 10) : cc;
    std::vector<int64_t> w_r0, w_rec;
    std::vector<int32_t> w_split;
    std::vector<int64_t> r_rec(static_cast<size_t>(nn));
    std::vector<int32_t> r_stride(static_cast<size_t>(nn));
    int64_t rec_total = 0, steps_total = 0, lds_max = 0;
    // Small meshes: at one row per lane the launch would hold fewer than
    // min_waves waves (default 2048: 8 per CU on 256 CUs); start from a
    // larger split so that the S lanes of a row share its incidences instead.
    const int64_t min_waves = P->min_waves;
    int s_min = 1;
    while (s_min * 2 <= 64 / unit && (nn * unit * s_min + 63) / 64 < min_waves) s_min <<= 1;
    P->min_split = s_min;
    int64_t r = 0;
    while (r < nn) {
        int chosen_R = 0, chosen_S = 0;
        for (int S = s_min; S <= 64 / unit; S <<= 1) {
            // power-of-two splits: the S parts of a row share its work
            int R = 64 / (unit * S);
            if (R < 1) break;
            const int64_t r1 = std::min<int64_t>(r + R, nn);
            const int64_t bytes = (hptr[size_t(r1)] - hptr[size_t(r)]) * cc * 8;
            if (bytes <= P->lds_budget || R == 1) {
                chosen_R = int(r1 - r);
                chosen_S = S;
                if (bytes <= P->lds_budget) break;
            }
        }
        const int64_t r1 = r + chosen_R;
        const int64_t bytes = (hptr[size_t(r1)] - hptr[size_t(r)]) * cc * 8;
        if (bytes > kMaxWaveLds) {
            set_error("the CSR span of one node row exceeds 64 KiB of LDS");
            return kCapacity;
        }
        lds_max = std::max(lds_max, bytes);
        int64_t steps = 0;
        for (int64_t q = r; q < r1; ++q) steps = std::max(steps, hinc[size_t(q + 1)] - hinc[size_t(q)]);
        // incidence layers padded to a multiple of S: the S lanes of a row
        // may take layers st*S + p (incidence split)
        steps = (steps + chosen_S - 1) / chosen_S * chosen_S;
        w_r0.push_back(r);
        w_rec.push_back(rec_total);
        w_split.push_back(chosen_S);
        for (int64_t q = r; q < r1; ++q) {
            r_rec[size_t(q)] = rec_total + (q - r);
            r_stride[size_t(q)] = chosen_R;
        }
        rec_total += steps * chosen_R;
        steps_total += steps;
        r = r1;
    }
    w_r0.push_back(nn);
    w_rec.push_back(rec_total);
    // XCD segments: 8 contiguous wave ranges of about equal record count
    {
        const int64_t nw = int64_t(w_split.size());
        int64_t x = 0;
        P->xcd_seg_host[0] = 0;
        for (int64_t w = 0; w < nw && x < 7; ++w) {
            if (w_rec[size_t(w + 1)] * 8 >= rec_total * (x + 1)) P->xcd_seg_host[++x] = w + 1;
        }
        while (x < 8) P->xcd_seg_host[++x] = nw;
        P->xcd_seg_host[8] = nw;
        for (int k = 1; k <= 8; ++k)
            if (P->xcd_seg_host[k] < P->xcd_seg_host[k - 1]) P->xcd_seg_host[k] = P->xcd_seg_host[k - 1];
    }
    P->nwaves = int64_t(w_split.size());
    P->total_records = rec_total;
    P->total_steps = steps_total;
    P->lds_bytes = int32_t(lds_max);
    if (int rc = P->xcd_seg.alloc(9)) return rc;
    SKFA_HIP(hipMemcpyAsync(P->xcd_seg.p, P->xcd_seg_host, sizeof(int64_t) * 9, hipMemcpyHostToDevice, s));
    if (int rc = P->wave_r0.alloc(w_r0.size())) return rc;
    if (int rc = P->wave_rec.alloc(w_rec.size())) return rc;
    if (int rc = P->wave_split.alloc(std::max<size_t>(w_split.size(), 1))) return rc;
    std::vector<WaveMeta> w_meta(w_split.size());
    for (size_t w = 0; w < w_split.size(); ++w) {
        WaveMeta& m = w_meta[w];
        const int64_t R = w_r0[w + 1] - w_r0[w];
        const int64_t steps = (w_rec[w + 1] - w_rec[w]) / R;
        if (R < 1 || R > 64 || steps > 65535) {
            set_error("gather wave layout out of range");
            return kCapacity;
        }
        m.r0 = w_r0[w];
        m.rec0 = w_rec[w];
        m.lo = hptr[size_t(w_r0[w])];
        m.span = int32_t(hptr[size_t(w_r0[w + 1])] - m.lo);
        m.steps = uint16_t(steps);
        m.R = uint8_t(R);
        m.S = uint8_t(w_split[w]);
    }
    if (int rc = P->wave_meta.alloc(std::max<size_t>(w_meta.size(), 1))) return rc;
    if (!w_meta.empty())
        SKFA_HIP(hipMemcpyAsync(P->wave_meta.p, w_meta.data(), sizeof(WaveMeta) * w_meta.size(),
                                hipMemcpyHostToDevice, s));
    if (int rc = P->row_rec.alloc(std::max<int64_t>(nn, 1))) return rc;
    if (int rc = P->row_stride.alloc(std::max<int64_t>(nn, 1))) return rc;
    SKFA_HIP(hipMemcpyAsync(P->wave_r0.p, w_r0.data(), sizeof(int64_t) * w_r0.size(), hipMemcpyHostToDevice, s));
    SKFA_HIP(hipMemcpyAsync(P->wave_rec.p, w_rec.data(), sizeof(int64_t) * w_rec.size(), hipMemcpyHostToDevice, s));
    if (!w_split.empty())
        SKFA_HIP(hipMemcpyAsync(P->wave_split.p, w_split.data(), sizeof(int32_t) * w_split.size(),
                                hipMemcpyHostToDevice, s));
    if (nn) {
        SKFA_HIP(hipMemcpyAsync(P->row_rec.p, r_rec.data(), sizeof(int64_t) * nn, hipMemcpyHostToDevice, s));
        SKFA_HIP(hipMemcpyAsync(P->row_stride.p, r_stride.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, s));
    }
    P->off_bytes = P->max_row_len <= 256 ? 1 : 2;
    if (P->max_row_len > 65536) {
        set_error("a node row has more than 65536 columns");
        return kCapacity;
    }
    const int raw = 4 + n * P->off_bytes;
    P->rec_bytes = raw <= 8 ? 8 : raw <= 16 ? 16 : raw <= 32 ? 32 : 64;
    const size_t rec_bytes_total = size_t(rec_total) * P->rec_bytes;
    if (int rc = P->records.alloc(rec_bytes_total ? rec_bytes_total : 1)) return rc;
    SKFA_HIP(hipMemsetAsync(P->records.p, 0xFF, rec_bytes_total ? rec_bytes_total : 1, s));
    k_row_analyze<<<rows_grid, 64, 0, s>>>(conn, n, P->inc_ptr.p, P->inc.p, nn, 1, nullptr, P->indptr.p,
                                           P->indices.p, P->row_rec.p, P->row_stride.p, P->records.p,
                                           P->rec_bytes, P->off_bytes, status.p);
    SKFA_LAUNCH_CHECK();
    SKFA_HIP(hipStreamSynchronize(s));
    return kOk;
}

// ------------------------------------------------------------------ cross plans
__global__ void k_check_range(const int32_t* __restrict__ conn, int64_t total, int64_t nnodes,
                              int32_t* __restrict__ status) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x) {
        const int32_t v = conn[i];
        if (v < 0 || v >= nnodes) atomicOr(status, 1);
    }
}

// Rectangular pattern: incidences of the row field's nodes, candidates from
// the column field's connectivity, one record per incidence.
static int cross_plan_build(skfa_cross_plan* P, const int32_t* rconn, const int32_t* cconn, hipStream_t s) {
    const int64_t E = P->E, nn = P->row_nodes;
    const int nr = P->n_row, nc = P->n_col;
    DevBuf<int32_t> cnt, status;
    if (int rc = cnt.alloc(nn + 1)) return rc;
    if (int rc = status.alloc(2)) return rc;
    SKFA_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (nn + 1), s));
    SKFA_HIP(hipMemsetAsync(status.p, 0, sizeof(int32_t) * 2, s));
    k_inc_count<<<grid_for(E * nr, 256), 256, 0, s>>>(rconn, E * nr, nn, cnt.p, status.p);
    k_check_range<<<grid_for(E * nc, 256), 256, 0, s>>>(cconn, E * nc, P->col_nodes, status.p);
    SKFA_LAUNCH_CHECK();
    if (int rc = P->inc_ptr.alloc(nn + 1)) return rc;
    if (int rc = exclusive_scan_i32_to_i64(cnt.p, P->inc_ptr.p, nn, s)) return rc;
    int32_t st = 0;
    SKFA_HIP(hipMemcpyAsync(&st, status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    if (st & 1) {
        set_error("element node index is out of bounds");
        return kInvalid;
    }
    if (int rc = P->inc.alloc(size_t(E * nr))) return rc;
    SKFA_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (nn + 1), s));
    k_inc_fill<<<grid_for(E * nr, 256), 256, 0, s>>>(rconn, E, nr, nn, P->inc_ptr.p, cnt.p, P->inc.p);
    SKFA_LAUNCH_CHECK();
    k_inc_sort<<<grid_for(nn, 256), 256, 0, s>>>(P->inc_ptr.p, P->inc.p, nn);
    SKFA_LAUNCH_CHECK();

    DevBuf<int32_t> rowlen;
    if (int rc = rowlen.alloc(nn + 1)) return rc;
    const unsigned rows_grid = unsigned(std::min<int64_t>(std::max<int64_t>(nn, 1), 1 << 20));
    k_row_analyze<<<rows_grid, 64, 0, s>>>(cconn, nc, P->inc_ptr.p, P->inc.p, nn, 0, rowlen.p, nullptr, nullptr,
                                           nullptr, nullptr, nullptr, 0, 1, status.p);
    SKFA_LAUNCH_CHECK();
    SKFA_HIP(hipMemcpyAsync(&st, status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    if (st & 2) {
        set_error("a node row touches more than 4096 element columns");
        return kCapacity;
    }
    if (int rc = P->indptr.alloc(nn + 1)) return rc;
    if (int rc = exclusive_scan_i32_to_i64(rowlen.p, P->indptr.p, nn, s)) return rc;
    SKFA_HIP(hipMemcpyAsync(&P->nnz_node, P->indptr.p + nn, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipMemsetAsync(status.p, 0, sizeof(int32_t) * 2, s));
    k_max_diff_i64<<<grid_for(nn, 256, 1024), 256, 0, s>>>(P->indptr.p, nn, status.p);
    k_max_diff_i64<<<grid_for(nn, 256, 1024), 256, 0, s>>>(P->inc_ptr.p, nn, status.p + 1);
    SKFA_LAUNCH_CHECK();
    int32_t mx[2];
    SKFA_HIP(hipMemcpyAsync(mx, status.p, sizeof(mx), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    P->max_row_len = mx[0];
    P->max_inc = mx[1];
    if (P->nnz() > INT32_MAX) {
        set_error("rectangular pattern exceeds 2^31 nonzeros");
        return kCapacity;
    }
    if (P->max_row_len > 65536) {
        set_error("a node row has more than 65536 columns");
        return kCapacity;
    }
    if (int rc = P->indices.alloc(size_t(P->nnz_node))) return rc;
    P->off_bytes = P->max_row_len <= 256 ? 1 : 2;
    const int raw = 4 + nc * P->off_bytes;
    P->rec_bytes = raw <= 8 ? 8 : raw <= 16 ? 16 : raw <= 32 ? 32 : 64;
    const size_t rec_total = size_t(E * nr) * P->rec_bytes;
    if (int rc = P->records.alloc(rec_total ? rec_total : 1)) return rc;
    k_row_analyze<<<rows_grid, 64, 0, s>>>(cconn, nc, P->inc_ptr.p, P->inc.p, nn, 1, nullptr, P->indptr.p,
                                           P->indices.p, P->inc_ptr.p, nullptr, P->records.p, P->rec_bytes,
                                           P->off_bytes, status.p);
    SKFA_LAUNCH_CHECK();
    SKFA_HIP(hipStreamSynchronize(s));
    return kOk;
}

// ------------------------------------------------------------------ CSR export
__global__ void k_export_csr(const int64_t* __restrict__ nptr, const int32_t* __restrict__ nidx, int64_t nnodes, int C,
                             int CC, int64_t* __restrict__ indptr, int32_t* __restrict__ indices) {
    // row C*n+c has columns CC*m+d for each m in node row n, m-major d-minor
    // (C row components, CC column components)
    for (int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < nnodes * C;
         r += int64_t(gridDim.x) * blockDim.x) {
        const int64_t node = r / C;
        const int c = int(r - node * C);
        const int64_t len = nptr[node + 1] - nptr[node];
        const int64_t start = int64_t(C) * CC * nptr[node] + int64_t(c) * CC * len;
        indptr[r] = start;
        for (int64_t j = 0; j < len; ++j) {
            const int32_t m = nidx[nptr[node] + j];
            for (int d = 0; d < CC; ++d) indices[start + j * CC + d] = m * CC + d;
        }
        if (r == nnodes * C - 1) indptr[r + 1] = start + len * CC;
    }
}

// ------------------------------------------------------------------ edges
// Open-addressing table of (lo,hi) keys; slot keeps the smallest item id
// (item = element*6 + local edge) that carries the key.
constexpr unsigned long long kEmpty = ~0ull;

__device__ inline unsigned long long edge_hash(unsigned long long key) {
    key ^= key >> 33;
    key *= 0xff51afd7ed558ccdull;
    key ^= key >> 33;
    key *= 0xc4ceb9fe1a85ec53ull;
    key ^= key >> 33;
    return key;
}

__constant__ int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

__global__ void k_edge_insert(const int32_t* __restrict__ vconn, int64_t E, int64_t nvert,
                              unsigned long long* __restrict__ keys, unsigned long long* __restrict__ minitem,
                              int64_t cap, int64_t* __restrict__ slot_of, unsigned long long* __restrict__ bad_cell,
                              int32_t* __restrict__ range_error) {
    const int64_t total = E * 6;
    for (int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; it < total; it += int64_t(gridDim.x) * blockDim.x) {
        const int64_t e = it / 6;
        const int l = int(it - e * 6);
        const int32_t* c = vconn + e * 4;
        if (l == 0) {
            bool bad = false;
            for (int i = 0; i < 4; ++i) {
                if (c[i] < 0 || c[i] >= nvert) atomicOr(range_error, 1);
                for (int j = i + 1; j < 4; ++j)
                    if (c[i] == c[j]) bad = true;
            }
            if (bad) atomicMin(bad_cell, (unsigned long long)e);
        }
        const int64_t f = c[kTetEdge[l][0]], s = c[kTetEdge[l][1]];
        if (f < 0 || f >= nvert || s < 0 || s >= nvert || f == s) {
            slot_of[it] = -1;
            continue;
        }
        const unsigned long long lo = f < s ? f : s, hi = f < s ? s : f;
        const unsigned long long key = lo * (unsigned long long)nvert + hi;
        unsigned long long slot = edge_hash(key) & (unsigned long long)(cap - 1);
        while (true) {
            const unsigned long long prev = atomicCAS(&keys[slot], kEmpty, key);
            if (prev == kEmpty || prev == key) break;
            slot = (slot + 1) & (unsigned long long)(cap - 1);
        }
        atomicMin(&minitem[slot], (unsigned long long)it);
        slot_of[it] = int64_t(slot);
    }
}

__global__ void k_edge_first(int64_t total, const int64_t* __restrict__ slot_of,
                             const unsigned long long* __restrict__ minitem, int32_t* __restrict__ flag) {
    for (int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; it < total; it += int64_t(gridDim.x) * blockDim.x) {
        const int64_t sl = slot_of[it];
        flag[it] = (sl >= 0 && minitem[sl] == (unsigned long long)it) ? 1 : 0;
    }
}

__global__ void k_edge_assign(const int32_t* __restrict__ vconn, int64_t E, const int64_t* __restrict__ slot_of,
                              const unsigned long long* __restrict__ minitem, const int64_t* __restrict__ rank,
                              int32_t* __restrict__ elem_edges, int8_t* __restrict__ signs,
                              int64_t* __restrict__ edge_lo, int64_t* __restrict__ edge_hi) {
    const int64_t total = E * 6;
    for (int64_t it = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; it < total; it += int64_t(gridDim.x) * blockDim.x) {
        const int64_t sl = slot_of[it];
        if (sl < 0) continue;
        const int64_t e = it / 6;
        const int l = int(it - e * 6);
        const int32_t* c = vconn + e * 4;
        const int32_t f = c[kTetEdge[l][0]], s = c[kTetEdge[l][1]];
        const int64_t first = int64_t(minitem[sl]);
        const int64_t id = rank[first];
        elem_edges[it] = int32_t(id);
        signs[it] = f < s ? 1 : -1;
        if (first == it && edge_lo) {
            edge_lo[id] = f < s ? f : s;
            edge_hi[id] = f < s ? s : f;
        }
    }
}

}  // namespace skfa

using namespace skfa;

extern "C" {

int skfa_plan_create(const int32_t* node_conn, int64_t E, int32_t n, int64_t nnodes, int32_t components,
                     void* stream, skfa_plan** out) {
    SKFA_REQUIRE(out != nullptr, "plan output pointer is null");
    *out = nullptr;
    SKFA_REQUIRE(E >= 0 && nnodes >= 0, "negative sizes");
    SKFA_REQUIRE(n >= 1 && n <= kMaxLocal, "nodes per element must be in [1, 16]");
    SKFA_REQUIRE(E < kMaxElements, "element count exceeds 2^27");
    SKFA_REQUIRE(components == 1 || components == 2 || components == 3, "components must be 1, 2 or 3");
    SKFA_REQUIRE(E == 0 || node_conn != nullptr, "connectivity pointer is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    skfa_plan* P = new skfa_plan();
    P->E = E;
    P->n = n;
    P->nnodes = nnodes;
    P->C = components;
    int rc = plan_build(P, node_conn, s);
    if (rc != kOk) {
        delete P;
        return rc;
    }
    *out = P;
    return kOk;
}

int skfa_plan_set_tuning(skfa_plan* P, int32_t key, int32_t value) {
    SKFA_REQUIRE(P != nullptr, "plan is null");
    switch (key) {
        case SKFA_TUNE_XCD_CHUNK:
            SKFA_REQUIRE(value >= 0 && value <= 4096, "xcd chunk out of range");
            P->tune_xcd_chunk = value;
            return kOk;
        case SKFA_TUNE_PLACEMENT:
            SKFA_REQUIRE(value >= 0 && value <= 2, "placement must be 0, 1 or 2");
            P->tune_placement = value;
            return kOk;
        case SKFA_TUNE_STREAMING:
            SKFA_REQUIRE(value == 0 || value == 1, "streaming flag must be 0 or 1");
            P->tune_streaming = value;
            return kOk;
        case SKFA_TUNE_LDS_ATOMIC:
            // the gathers always accumulate with ds_add_f64; 1 is kept as
            // the accepted value of this key
            SKFA_REQUIRE(value == 1, "LDS accumulation is always ds_add_f64: lds atomic flag must be 1");
            return kOk;
        case SKFA_TUNE_ELEMENT_TABLE:
            SKFA_REQUIRE(value >= -1 && value <= 1, "element table flag must be -1 (auto), 0 or 1");
            P->tune_table = value;
            return kOk;
        case SKFA_TUNE_MFMA:
            SKFA_REQUIRE(value >= -1 && value <= 2, "MFMA mode must be -1 (auto), 0, 1 or 2");
            P->tune_mfma = value;
            return kOk;
        case SKFA_TUNE_ROW_SPLIT:
            SKFA_REQUIRE(value >= 0 && value <= 3,
                         "row split must be 0 (incidence), 1 (entry), 2 (column) or 3 (concurrent incidence)");
            P->tune_split = value;
            return kOk;
        case SKFA_TUNE_ROW_FORM:
            SKFA_REQUIRE(value >= -1 && value <= 1, "row form must be -1 (auto), 0 (table) or 1 (barycentric)");
            P->tune_kform = value;
            return kOk;
        default:
            set_error("unknown tuning key");
            return kInvalid;
    }
}

int skfa_plan_create_ex(const int32_t* node_conn, int64_t E, int32_t n, int64_t nnodes, int32_t components,
                        int32_t lds_budget, int64_t min_waves, void* stream, skfa_plan** out) {
    return skfa_plan_create_layout(node_conn, E, n, nnodes, components, lds_budget, min_waves, 0, stream, out);
}

int skfa_plan_create_layout(const int32_t* node_conn, int64_t E, int32_t n, int64_t nnodes, int32_t components,
                            int32_t lds_budget, int64_t min_waves, int32_t row_lanes, void* stream,
                            skfa_plan** out) {
    SKFA_REQUIRE(out != nullptr, "plan output pointer is null");
    SKFA_REQUIRE(row_lanes == 0 || row_lanes == 4, "row lanes must be 0 (default layout) or 4");
    SKFA_REQUIRE(row_lanes == 0 || (components == 3 && n == 10),
                 "the 4-lane row layout is for 10-node 3-component fields");
    *out = nullptr;
    SKFA_REQUIRE(lds_budget >= 2048 && lds_budget <= int32_t(kMaxWaveLds), "LDS budget out of range");
    SKFA_REQUIRE(E >= 0 && nnodes >= 0, "negative sizes");
    SKFA_REQUIRE(min_waves >= 0, "minimum wave count must be non-negative");
    SKFA_REQUIRE(n >= 1 && n <= kMaxLocal, "nodes per element must be in [1, 16]");
    SKFA_REQUIRE(E < kMaxElements, "element count exceeds 2^27");
    SKFA_REQUIRE(components == 1 || components == 2 || components == 3, "components must be 1, 2 or 3");
    SKFA_REQUIRE(E == 0 || node_conn != nullptr, "connectivity pointer is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    skfa_plan* P = new skfa_plan();
    P->E = E;
    P->n = n;
    P->nnodes = nnodes;
    P->C = components;
    P->lds_budget = lds_budget;
    P->min_waves = min_waves;
    P->row_lanes = row_lanes;
    int rc = plan_build(P, node_conn, s);
    if (rc != kOk) {
        delete P;
        return rc;
    }
    *out = P;
    return kOk;
}

int skfa_plan_create_budget(const int32_t* node_conn, int64_t E, int32_t n, int64_t nnodes, int32_t components,
                            int32_t lds_budget, void* stream, skfa_plan** out) {
    return skfa_plan_create_ex(node_conn, E, n, nnodes, components, lds_budget, 2048, stream, out);
}

int skfa_plan_destroy(skfa_plan* plan) {
    delete plan;
    return kOk;
}

int skfa_plan_info(const skfa_plan* P, int64_t* info) {
    SKFA_REQUIRE(P != nullptr && info != nullptr, "null argument");
    info[0] = P->nrows();
    info[1] = P->nnz();
    info[2] = P->nnz_node;
    info[3] = P->max_row_len;
    info[4] = P->max_inc;
    info[5] = P->nwaves;
    info[6] = P->total_records;
    info[7] = P->rec_bytes;
    info[8] = P->lds_bytes;
    return kOk;
}

int skfa_plan_export_csr(const skfa_plan* P, int64_t* indptr, int32_t* indices, void* stream) {
    SKFA_REQUIRE(P != nullptr && indptr != nullptr, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->nnodes == 0) {
        SKFA_HIP(hipMemsetAsync(indptr, 0, sizeof(int64_t), s));
        return kOk;
    }
    k_export_csr<<<grid_for(P->nnodes * P->C, 256), 256, 0, s>>>(P->indptr.p, P->indices.p, P->nnodes, P->C, P->C,
                                                                  indptr, indices);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_plan_export_map(const skfa_plan* P, int64_t* inc_ptr, int32_t* inc, void* stream) {
    SKFA_REQUIRE(P != nullptr && inc_ptr != nullptr && inc != nullptr, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    SKFA_HIP(hipMemcpyAsync(inc_ptr, P->inc_ptr.p, sizeof(int64_t) * (P->nnodes + 1), hipMemcpyDeviceToDevice, s));
    if (P->E * P->n)
        SKFA_HIP(hipMemcpyAsync(inc, P->inc.p, sizeof(int32_t) * P->E * P->n, hipMemcpyDeviceToDevice, s));
    return kOk;
}

int skfa_cross_plan_create(const int32_t* row_conn, int32_t row_nodes_per_element, int64_t row_nodes,
                           int32_t row_components, const int32_t* col_conn, int32_t col_nodes_per_element,
                           int64_t col_nodes, int32_t col_components, int64_t E, void* stream,
                           skfa_cross_plan** out) {
    SKFA_REQUIRE(out != nullptr, "plan output pointer is null");
    *out = nullptr;
    SKFA_REQUIRE(E >= 0 && row_nodes >= 0 && col_nodes >= 0, "negative sizes");
    SKFA_REQUIRE(row_nodes_per_element >= 1 && row_nodes_per_element <= kMaxLocal &&
                     col_nodes_per_element >= 1 && col_nodes_per_element <= kMaxLocal,
                 "nodes per element must be in [1, 16]");
    SKFA_REQUIRE(E < kMaxElements, "element count exceeds 2^27");
    SKFA_REQUIRE(row_components >= 1 && row_components <= 3 && col_components >= 1 && col_components <= 3,
                 "components must be 1, 2 or 3");
    SKFA_REQUIRE(E == 0 || (row_conn != nullptr && col_conn != nullptr), "connectivity pointer is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    skfa_cross_plan* P = new skfa_cross_plan();
    P->E = E;
    P->n_row = row_nodes_per_element;
    P->n_col = col_nodes_per_element;
    P->row_nodes = row_nodes;
    P->col_nodes = col_nodes;
    P->C_row = row_components;
    P->C_col = col_components;
    int rc = cross_plan_build(P, row_conn, col_conn, s);
    if (rc != kOk) {
        delete P;
        return rc;
    }
    *out = P;
    return kOk;
}

int skfa_cross_plan_destroy(skfa_cross_plan* plan) {
    delete plan;
    return kOk;
}

int skfa_cross_plan_info(const skfa_cross_plan* P, int64_t* info) {
    SKFA_REQUIRE(P != nullptr && info != nullptr, "null argument");
    info[0] = P->nrows();
    info[1] = P->ncols();
    info[2] = P->nnz();
    info[3] = P->nnz_node;
    info[4] = P->max_row_len;
    info[5] = P->max_inc;
    return kOk;
}

int skfa_cross_plan_export_csr(const skfa_cross_plan* P, int64_t* indptr, int32_t* indices, void* stream) {
    SKFA_REQUIRE(P != nullptr && indptr != nullptr, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (P->row_nodes == 0) {
        SKFA_HIP(hipMemsetAsync(indptr, 0, sizeof(int64_t), s));
        return kOk;
    }
    k_export_csr<<<grid_for(P->row_nodes * P->C_row, 256), 256, 0, s>>>(P->indptr.p, P->indices.p, P->row_nodes,
                                                                         P->C_row, P->C_col, indptr, indices);
    SKFA_LAUNCH_CHECK();
    return kOk;
}

int skfa_edge_topology(const int32_t* vconn, int64_t E, int64_t nvertices, int32_t* elem_edges, int8_t* signs,
                       int64_t* edge_lo, int64_t* edge_hi, int64_t* nedges, int64_t* bad_cell, void* stream) {
    SKFA_REQUIRE(nedges != nullptr && bad_cell != nullptr, "null argument");
    SKFA_REQUIRE(E >= 0 && nvertices >= 0, "negative sizes");
    SKFA_REQUIRE(E * 6 < (int64_t(1) << 31), "element count too large for edge numbering");
    hipStream_t s = static_cast<hipStream_t>(stream);
    *nedges = 0;
    *bad_cell = -1;
    if (E == 0) return kOk;
    const int64_t total = E * 6;
    int64_t cap = 1;
    while (cap < 2 * total) cap <<= 1;
    DevBuf<unsigned long long> keys, minitem;
    DevBuf<int64_t> slot_of, rank;
    DevBuf<int32_t> flag;
    DevBuf<unsigned long long> status;
    DevBuf<int32_t> range;
    if (int rc = range.alloc(1)) return rc;
    if (int rc = keys.alloc(cap)) return rc;
    if (int rc = minitem.alloc(cap)) return rc;
    if (int rc = slot_of.alloc(total)) return rc;
    if (int rc = rank.alloc(total + 1)) return rc;
    if (int rc = flag.alloc(total)) return rc;
    if (int rc = status.alloc(1)) return rc;
    SKFA_HIP(hipMemsetAsync(keys.p, 0xFF, sizeof(unsigned long long) * cap, s));
    SKFA_HIP(hipMemsetAsync(minitem.p, 0xFF, sizeof(unsigned long long) * cap, s));
    SKFA_HIP(hipMemsetAsync(status.p, 0xFF, sizeof(unsigned long long), s));
    SKFA_HIP(hipMemsetAsync(range.p, 0, sizeof(int32_t), s));
    k_edge_insert<<<grid_for(total, 256), 256, 0, s>>>(vconn, E, nvertices, keys.p, minitem.p, cap, slot_of.p,
                                                        status.p, range.p);
    SKFA_LAUNCH_CHECK();
    unsigned long long bad = 0;
    int32_t range_bad = 0;
    SKFA_HIP(hipMemcpyAsync(&bad, status.p, sizeof(bad), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipMemcpyAsync(&range_bad, range.p, sizeof(range_bad), hipMemcpyDeviceToHost, s));
    SKFA_HIP(hipStreamSynchronize(s));
    if (range_bad) {
        set_error("connectivity index is out of bounds");
        return kInvalid;
    }
    if (bad != ~0ull) {
        *bad_cell = int64_t(bad);
        set_error("cell " + std::to_string(bad) + " repeats a corner vertex");
        return kInvalid;
    }
    k_edge_first<<<grid_for(total, 256), 256, 0, s>>>(total, slot_of.p, minitem.p, flag.p);
    SKFA_LAUNCH_CHECK();
    if (int rc = exclusive_scan_i32_to_i64(flag.p, rank.p, total, s)) return rc;
    SKFA_HIP(hipMemcpyAsync(nedges, rank.p + total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    k_edge_assign<<<grid_for(total, 256), 256, 0, s>>>(vconn, E, slot_of.p, minitem.p, rank.p, elem_edges, signs,
                                                        edge_lo, edge_hi);
    SKFA_LAUNCH_CHECK();
    SKFA_HIP(hipStreamSynchronize(s));
    return kOk;
}

}  // extern "C"
