// Element->CSR gather plan: the device-resident replacement for the CSR
// pattern, element->CSR scatter map and element colouring that the
// reference builds in BilinearFormAssembler's constructor
// (cpp/src/assembly/bilinear_form_assembler.cpp:18-38,111-140) and in
// build_edge_csr_pattern (cpp/src/assembly/edge_topology.cpp:86-154).
#pragma once

#include <vector>

#include "skfa_internal.h"

struct skfa_plan {
    int64_t E = 0;        // elements
    int32_t n = 0;        // nodes per element (local rows of the node-level map)
    int64_t nnodes = 0;   // node-level rows
    int32_t C = 1;        // field components (DOF = node * C + c)
    int32_t lanes_per_row = 1;  // C*C lanes cooperate on one node row
    int32_t rows_per_wave = 64; // node rows handled by one 64-lane wave
    int32_t max_row_len = 0;    // longest node row
    int32_t max_inc = 0;        // largest element count around one node

    // node-level CSR pattern
    skfa::DevBuf<int64_t> indptr;  // nnodes+1
    skfa::DevBuf<int32_t> indices; // nnz_node
    int64_t nnz_node = 0;

    // node -> (element, local node) incidence, sorted by element
    skfa::DevBuf<int64_t> inc_ptr; // nnodes+1
    skfa::DevBuf<int32_t> inc;     // E*n, packed (e << 4) | a

    // wave-lockstep gather records: record (wave_ptr[w]+s)*R + i holds the
    // s-th incidence of node row w*R+i and the in-row offsets of the
    // element's n columns.  Padding records carry packed = -1.
    int64_t nwaves = 0;
    skfa::DevBuf<int64_t> wave_ptr; // nwaves+1
    int32_t rec_bytes = 0;
    int32_t off_bytes = 1;
    int64_t total_steps = 0;
    skfa::DevBuf<uint8_t> records;

    // waves grouped by the LDS footprint of their CSR span
    int32_t nbuckets = 0;
    int64_t bucket_lds[8] = {0};   // bytes of LDS per wave in bucket
    int64_t bucket_start[9] = {0}; // offsets into bucket_waves
    skfa::DevBuf<int32_t> bucket_waves;

    int64_t nrows() const { return nnodes * C; }
    int64_t nnz() const { return nnz_node * C * C; }
};

namespace skfa {
constexpr int kMaxCandidates = 4096;   // per node row: incidences * n
constexpr int kMaxLocal = 16;          // local nodes per element (packed in 4 bits)
constexpr int64_t kMaxElements = (int64_t(1) << 27);
constexpr int64_t kMaxWaveLds = 64 * 1024;  // bytes of LDS for one wave's CSR span
}  // namespace skfa
