// Element->CSR gather plan: the device-resident replacement for the CSR
// pattern, element->CSR scatter map and element colouring that the
// reference builds in BilinearFormAssembler's constructor
// (cpp/src/assembly/bilinear_form_assembler.cpp:18-38,111-140) and in
// build_edge_csr_pattern (cpp/src/assembly/edge_topology.cpp:86-154).
#pragma once

#include <vector>

#include "skfa_internal.h"

namespace skfa {
// Per-wave constants of the gather, one 32-byte scalar load per wave.
struct WaveMeta {
    int64_t r0;       // first node row
    int64_t rec0;     // first gather record
    int64_t lo;       // node-level CSR offset of the first row
    int32_t span;     // node-level CSR entries of the wave's rows
    uint16_t steps;   // incidence layers (a multiple of S)
    uint8_t R, S;     // rows, lanes (parts) per row
};
static_assert(sizeof(WaveMeta) == 32, "one s_load_dwordx8 per wave");
}  // namespace skfa

struct skfa_plan {
    int64_t E = 0;        // elements
    int32_t n = 0;        // nodes per element (local rows of the node-level map)
    int64_t nnodes = 0;   // node-level rows
    int32_t C = 1;        // field components (DOF = node * C + c)
    int32_t max_row_len = 0;    // longest node row
    int32_t max_inc = 0;        // largest element count around one node

    // node-level CSR pattern
    skfa::DevBuf<int64_t> indptr;  // nnodes+1
    skfa::DevBuf<int32_t> indices; // nnz_node
    int64_t nnz_node = 0;

    // node -> (element, local node) incidence, sorted by element
    skfa::DevBuf<int64_t> inc_ptr; // nnodes+1
    skfa::DevBuf<int32_t> inc;     // E*n, packed (e << 4) | a

    // Wave-lockstep gather records.  Wave w owns node rows
    // [wave_r0[w], wave_r0[w+1]) (R_w = their count) and split factor S_w:
    // S_w parts (lane groups) share one node row; with the default incidence
    // split part p takes the row's incidences p, p+S_w, ...  R_w and S_w are
    // chosen so the wave's CSR span fits the LDS budget.  Record
    // wave_rec[w] + s*R_w + i holds the s-th incidence of row wave_r0[w]+i
    // and the in-row offsets of the element's n columns; the layer count is
    // a multiple of S_w and padding records carry packed = -1.
    int64_t nwaves = 0;
    skfa::DevBuf<int64_t> wave_r0;    // nwaves+1
    skfa::DevBuf<int64_t> wave_rec;   // nwaves+1
    skfa::DevBuf<int32_t> wave_split; // nwaves
    skfa::DevBuf<skfa::WaveMeta> wave_meta;  // nwaves: the above, gathered per wave
    skfa::DevBuf<int64_t> row_rec;    // nnodes: record of the row's first incidence
    skfa::DevBuf<int32_t> row_stride; // nnodes: R of the row's wave
    int32_t rec_bytes = 0;
    int32_t off_bytes = 1;
    int64_t total_records = 0;
    int64_t total_steps = 0;          // sum over waves of steps (for reporting)
    int32_t lds_budget = 16 * 1024;   // bytes of LDS per wave requested at plan time
    int32_t lds_bytes = 0;            // largest span of any wave, bytes
    int32_t min_split = 1;            // smallest S of the layout (> 1 for small meshes)
    int64_t min_waves = 2048;         // small meshes: raise S until the launch has this many waves
    int32_t row_lanes = 0;            // lanes per node row and part: 0 = the kernel default (C*C, or 10 for
                                      // 10-node 3-component fields); 4 = the vector-P2 MFMA gather's 16-row waves
    skfa::DevBuf<uint8_t> records;

    // per-element payload of the element-constant fast path (allocated on
    // first use, E * stride doubles)
    skfa::DevBuf<double> payload;
    // |det J| per element for the constant-load linear form (first use)
    skfa::DevBuf<double> elem_det;
    // full-size values and load of skfa_assemble_bilinear_load with output
    // maps when the load is not fused into the gather (first use)
    skfa::DevBuf<double> full_values;
    skfa::DevBuf<double> full_load;
    // element tables of the symmetric scalar gather (FP64 and FP32 copies)
    skfa::DevBuf<double> ntab;
    skfa::DevBuf<float> ntab32;

    // launch tuning (skfa_plan_set_tuning): XCD chunk of waves (0 = plain
    // order) and LDS accumulation by ds_add (1) or read-add-write (0)
    int32_t tune_table = -1;      // element-table gather for symmetric scalar forms: -1 auto, 0 off, 1 on
    int32_t tune_split = 0;       // quadrature gather rows with S lanes: 0 incidence, 1 entry, 2 column split,
                                  // 3 concurrent incidence split
    int32_t tune_kform = -1;      // 10-node scalar rows: 1 barycentric structure, 0 full table, -1 auto (1)
    int32_t tune_mfma = -1;       // MFMA row contraction for scalar fields: -1/0 off, 1 all waves, 2 waves <= 16 rows
    int32_t tune_placement = 2;   // 0 plain, 1 XCD segments, 2 XCD chunks
    int32_t tune_xcd_chunk = 16;
    int32_t tune_streaming = 1;   // non-temporal record loads / value stores
    // contiguous wave segments of equal gather work, one per XCD
    int64_t xcd_seg_host[9] = {0};
    skfa::DevBuf<int64_t> xcd_seg;

    int64_t nrows() const { return nnodes * C; }
    int64_t nnz() const { return nnz_node * C * C; }
};

// Rectangular (row field x column field) pattern on one mesh: the device
// replacement for CrossBilinearAssembler::build_pattern
// (cpp/src/assembly/cross_bilinear_assembler.cpp:318-326).  Rows are the
// nodes of the row (test) field, columns the nodes of the column (trial)
// field; DOF = node * components + component on either side.
struct skfa_cross_plan {
    int64_t E = 0;
    int32_t n_row = 0, n_col = 0;        // nodes per element of each field
    int64_t row_nodes = 0, col_nodes = 0;
    int32_t C_row = 1, C_col = 1;
    int32_t max_row_len = 0, max_inc = 0;
    // row-node -> (element, local row node) incidences, sorted by element
    skfa::DevBuf<int64_t> inc_ptr;
    skfa::DevBuf<int32_t> inc;
    // node-level rectangular pattern
    skfa::DevBuf<int64_t> indptr;
    skfa::DevBuf<int32_t> indices;
    int64_t nnz_node = 0;
    // one record per incidence (record k = incidence k): packed (e << 4) | a,
    // then the in-row node offsets of the element's n_col column nodes
    skfa::DevBuf<uint8_t> records;
    int32_t rec_bytes = 0, off_bytes = 1;
    // cut-cell forms: per-cell node matrices (E, n_row, n_col) or node
    // vectors (E, n_row, C_row) FP64, allocated by the first cut assembly on
    // the plan and grown as needed (calls on one plan share it, so they must
    // not overlap)
    mutable skfa::DevBuf<double> cell_buffer;

    int64_t nrows() const { return row_nodes * C_row; }
    int64_t ncols() const { return col_nodes * C_col; }
    int64_t nnz() const { return nnz_node * C_row * C_col; }
};

namespace skfa {
constexpr int kMaxCandidates = 4096;   // per node row: incidences * n
constexpr int kMaxLocal = 16;          // local nodes per element (packed in 4 bits)
constexpr int64_t kMaxElements = (int64_t(1) << 27);
constexpr int64_t kMaxWaveLds = 64 * 1024;  // bytes of LDS for one wave's CSR span
}  // namespace skfa
