/*
 * skfem_amd.h -- C ABI of the MI355X (gfx950) element-assembly library
 * libskfem_amd.so.
 *
 * Drop-in boundary for the native half of scikit-fem-native's assembly path.
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository).  All array arguments are raw DEVICE pointers
 * (HIP), C-contiguous, unless marked "host".  `stream` is a hipStream_t
 * (NULL = default stream).  Every function returns 0 on success or a negative
 * status (-1 invalid argument, -2 HIP runtime error, -3 unsupported,
 * -4 documented capacity limit exceeded); skfa_last_error() returns the
 * calling thread's last message.  No function silently falls back to the CPU.
 *
 * Conventions shared with the reference:
 *   - DOF numbering dof(node, c) = node * C + c (python/skfemntv/basis.py:1768);
 *   - CSR rows sorted, unique columns, full element connectivity pattern kept
 *     (explicit zeros included), row = test function, column = trial function
 *     (cpp/src/assembly/bilinear_form_assembler.cpp:111-120, :101-102).
 */
#ifndef SKFEM_AMD_H
#define SKFEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKFA_ABI_VERSION 1

typedef struct skfa_plan skfa_plan;

const char* skfa_last_error(void);
int skfa_abi_version(void);
int skfa_device_name(char* buf, int32_t len);

/* ---------------------------------------------------------------------------
 * Gather plan = CSR pattern + element->CSR map for a continuous nodal space.
 * Replaces the constructor half of BilinearFormAssembler (pattern + int64
 * scatter map + greedy colouring,
 * cpp/src/assembly/bilinear_form_assembler.cpp:18-38,111-140), of
 * CrossBilinearAssembler for equal row/column spaces
 * (cpp/src/assembly/cross_bilinear_assembler.cpp:19-54,307-341) and
 * build_edge_csr_pattern (cpp/src/assembly/edge_topology.cpp:86-154).
 *   node_conn  (E, n) int32 node ids in [0, nnodes); the DOFs of element e are
 *              node_conn[e,a]*C + c (the reference's (E,n,C) dofs array).
 *   components C in {1,2,3}.
 * Limits: n <= 16, E < 2^27, <= 4096 element columns around one node, the
 * CSR span of one wave <= 64 KiB; exceeding them returns -4.
 * The plan owns its device memory (hipMalloc) until skfa_plan_destroy.
 */
int skfa_plan_create(const int32_t* node_conn, int64_t E, int32_t n, int64_t nnodes, int32_t components,
                     void* stream, skfa_plan** out);
/* Same with the LDS budget (bytes, 2 KiB..64 KiB) that sizes each wave's
 * CSR span: smaller budgets give more waves per CU, larger ones fewer
 * column-split lanes.  skfa_plan_create uses 16 KiB. */
int skfa_plan_create_budget(const int32_t* node_conn, int64_t E, int32_t n, int64_t nnodes, int32_t components,
                            int32_t lds_budget, void* stream, skfa_plan** out);
/* Same with the smallest launch size: while the gather would have fewer
 * than min_waves waves at one row per lane, the rows are split over more
 * lanes (skfa_plan_create_budget uses 2048; 0 keeps one row per lane at
 * every size, the layout of large meshes). */
int skfa_plan_create_ex(const int32_t* node_conn, int64_t E, int32_t n, int64_t nnodes, int32_t components,
                        int32_t lds_budget, int64_t min_waves, void* stream, skfa_plan** out);
/* Same with the lanes per node row of the gather waves: 0 = the default
 * layout; 4 (10-node 3-component fields only) = the layout of the vector-P2
 * MFMA gather, 16 rows per wave, each row one column of the MFMA tiles and
 * its 4 lanes the 4 slices of the contraction.  A plan with this layout
 * assembles affine elements with element-constant coefficients, a 4-point
 * reduced gradient rule and at most 4 mass points (else status -1).  Same
 * pattern as every other layout; no reference counterpart (a launch-shape
 * choice of this library, like the LDS budget). */
int skfa_plan_create_layout(const int32_t* node_conn, int64_t E, int32_t n, int64_t nnodes, int32_t components,
                            int32_t lds_budget, int64_t min_waves, int32_t row_lanes, void* stream,
                            skfa_plan** out);
int skfa_plan_destroy(skfa_plan* plan);
/* Launch tuning of an existing plan (results are bit-identical for every
 * setting of keys 1-4; key 5 selects between two formulations that differ
 * by rounding only).  Keys: */
#define SKFA_TUNE_XCD_CHUNK 1  /* waves per XCD chunk, 0 = plain dispatch order */
#define SKFA_TUNE_LDS_ATOMIC 2 /* LDS accumulation: only 1 (ds_add_f64, always used) is accepted */
#define SKFA_TUNE_PLACEMENT 3  /* 0 plain order, 1 one contiguous segment per XCD, 2 XCD chunks */
#define SKFA_TUNE_STREAMING 4  /* 1: non-temporal record loads and CSR value stores */
#define SKFA_TUNE_ELEMENT_TABLE 5 /* symmetric scalar forms: 1 element-table gather, 0 quadrature gather,
                                     -1 automatic (table for 4-node elements); rounding differs only */
#define SKFA_TUNE_ROW_SPLIT 6  /* quadrature gather, rows with S > 1 lanes: 0 incidence split (lane p takes
                                  the row's incidences p, p+S, ...; the S lanes add in turn; default),
                                  1 entry split (lane p computes entries p, p+S, ... of every element
                                  row), 2 column split (every lane computes the whole element row and
                                  adds the columns it owns); bit-identical results; 3 concurrent
                                  incidence split (as 0, but the S lanes add in the same
                                  instructions and the LDS atomic add serialises lanes that hit one
                                  address: rounding may differ from 0-2) */
#define SKFA_TUNE_MFMA 7       /* scalar fields, element-constant coefficients, <= 4-point reduced rules:
                                  element rows contracted on MFMA (16 rows x 4 points per instruction);
                                  0 off (default, -1 = default), 1 every wave, 2 waves of <= 16 rows
                                  only; rounding differs only */
#define SKFA_TUNE_ROW_FORM 8   /* 10-node gathers with SKFA_FLAG_BARYCENTRIC_TABLE: 1 scalar element rows
                                  from the barycentric structure (64 FMAs per row, not 120) and
                                  vector-P2 3x3 node blocks from the pair table (54 FMAs, not 108),
                                  0 from the quadrature tables, -1 automatic (1); rounding differs
                                  only */
int skfa_plan_set_tuning(skfa_plan* plan, int32_t key, int32_t value);
/* host info[9]: nrows, nnz, nnz_node, max_row_len, max_incidence, waves,
 * gather records, record bytes, LDS bytes of the largest wave span. */
int skfa_plan_info(const skfa_plan* plan, int64_t* info);
/* CSR pattern (indptr (nrows+1) int64, indices (nnz) int32): the reference's
 * BilinearFormAssembler.indptr / .indices (bilinear_form_assembler.cpp:75-76). */
int skfa_plan_export_csr(const skfa_plan* plan, int64_t* indptr, int32_t* indices, void* stream);
/* node -> (element<<4 | local node) incidence lists, sorted by element. */
int skfa_plan_export_map(const skfa_plan* plan, int64_t* inc_ptr, int32_t* inc, void* stream);

/* ---------------------------------------------------------------------------
 * Rectangular plan between two nodal fields on one mesh (row = test field,
 * column = trial field).  Replaces the constructor half of
 * CrossBilinearAssembler for different row and column spaces: pattern,
 * int64 scatter map and colouring
 * (cpp/src/assembly/cross_bilinear_assembler.cpp:19-54,318-341).
 *   row_conn (E, row_nodes_per_element) int32 row node ids in [0, row_nodes),
 *   col_conn (E, col_nodes_per_element) int32 column node ids in [0, col_nodes);
 *   row DOF = row node * row_components + c, column DOF likewise.
 * The pattern holds every (row DOF, column DOF) pair of every element,
 * sorted and unique per row, as the reference's build_pattern does.
 * Limits: nodes per element <= 16, E < 2^27, <= 4096 element columns
 * around one row node, nnz < 2^31. */
typedef struct skfa_cross_plan skfa_cross_plan;
int skfa_cross_plan_create(const int32_t* row_conn, int32_t row_nodes_per_element, int64_t row_nodes,
                           int32_t row_components, const int32_t* col_conn, int32_t col_nodes_per_element,
                           int64_t col_nodes, int32_t col_components, int64_t E, void* stream,
                           skfa_cross_plan** out);
int skfa_cross_plan_destroy(skfa_cross_plan* plan);
/* host info[6]: rows, columns, nnz, nnz_node, max_row_len (column nodes),
 * max incidences of one row node. */
int skfa_cross_plan_info(const skfa_cross_plan* plan, int64_t* info);
/* CSR pattern: indptr (rows+1) int64, indices (nnz) int32 (the reference's
 * CrossBilinearAssembler.indptr / .indices, cross_bilinear_assembler.cpp:203). */
int skfa_cross_plan_export_csr(const skfa_cross_plan* plan, int64_t* indptr, int32_t* indices, void* stream);

/* Cross assembly into CSR values (nnz FP64, overwritten).  Replaces
 * CrossBilinearAssembler::assemble (cpp/src/assembly/cross_bilinear_assembler.cpp:55-107,207-247,
 * contraction cpp/include/native_fem/cross_contraction.hpp:12-84) for two
 * tetrahedral nodal fields on one mesh, with the geometry computed on the
 * GPU (as skfa_assemble_bilinear does).
 *   tables   FP64 w[Q] | row phi[Q][nr] | row dphi[Q][nr][3] | col phi[Q][nc] |
 *            col dphi[Q][nc][3] | gdphi[Q][gnodes][3] (the last block is read
 *            only for isoparametric geometry, gconn != NULL)
 *   row_kind, col_kind: SKFA_CROSS_VALUE or SKFA_CROSS_GRADIENT
 *   mode     SKFA_CROSS_SCALAR: entry (a,r),(b,c) = [r == c] coef * (phi_a phi_b
 *            or grad phi_a . grad phi_b); needs equal component counts and
 *            row_kind == col_kind (the reference's scalar-coefficient branch);
 *            SKFA_CROSS_ROW_DIVERGENCE: dphi_a/dx_r * phi_b * coef (vector rows
 *            with 3 components, scalar columns: div(v) u);
 *            SKFA_CROSS_COLUMN_DIVERGENCE: phi_a * dphi_b/dx_c * coef (scalar rows,
 *            vector columns: div(u) v);
 *            SKFA_CROSS_TENSOR: sum_ij row_i(a) T[r,(i),c,(j)] col_j(b) with T the
 *            (E,Q,Cr,[3],Cc,[3]) tensor of cross_contraction.hpp:42-63.
 *   coefficient FP64 scalar (modes 0-2) or tensor block (mode 3) of element e,
 *            point q at coefficient + e*coef_stride_e + q*coef_stride_q
 *            (strides in doubles; 0 broadcasts); NULL = 1 (modes 0-2 only).
 *   rtab     optional reduced table rt[a][b][k][l] = sum_q w_q h_ak(q) h_bl(q)
 *            (nr x nc x [3] x [3]; h = phi for a value side, dphi/dxi for a
 *            gradient side).  With rtab, affine geometry (gconn NULL), modes
 *            0-2 and an element-constant coefficient (coef_stride_q 0) a
 *            reduced-table kernel runs for the field pairs (nodes, components)
 *            P1 x vector P2, vector P2 x P1, P1 x P2, P2 x P1, P1 x P1, P2 x P2;
 *            otherwise, or with rtab NULL, the quadrature kernel.
 * The CSR span of the rows of one wave is accumulated in LDS and every entry
 * adds its incidences in element order, so results are deterministic. */
#define SKFA_CROSS_VALUE 0
#define SKFA_CROSS_GRADIENT 1
#define SKFA_CROSS_SCALAR 0
#define SKFA_CROSS_ROW_DIVERGENCE 1
#define SKFA_CROSS_COLUMN_DIVERGENCE 2
#define SKFA_CROSS_TENSOR 3
int skfa_assemble_cross(const skfa_cross_plan* plan, const double* xyz, const int32_t* vconn,
                        const int32_t* gconn, int32_t gnodes, const double* tables, int32_t nq, const double* rtab,
                        int32_t row_kind, int32_t col_kind, int32_t mode, const double* coefficient,
                        int64_t coef_stride_e, int64_t coef_stride_q, double* values, void* stream);

/* ---------------------------------------------------------------------------
 * Bilinear assembly into CSR values (nnz FP64, overwritten).
 * Replaces BilinearFormAssembler::assemble (value / gradient /
 * symmetric_gradient / divergence coefficients,
 * cpp/src/assembly/bilinear_form_assembler.cpp:39-105) and the
 * gradient x gradient tensor path of CrossBilinearAssembler::assemble
 * (cpp/src/assembly/cross_bilinear_assembler.cpp:55-107,207-247 with
 * cpp/include/native_fem/cross_contraction.hpp:42-84), with the geometry of
 * tabulate_basis_geometry (cpp/src/assembly/basis_geometry.cpp:107-176)
 * computed on the GPU instead of read from (E,Q,n,3) tabulations.
 *   xyz      (npoints, 4) FP64 x,y,z,pad
 *   vconn    (E,4) int32 corner vertices (affine elements)
 *   gconn    (E,gnodes) int32 geometry nodes for isoparametric elements, or
 *            NULL for affine elements
 *   tables   FP64 w[Q] | phi[Q][n] | dphi[Q][n][3] | gphi[Q][gn] | gdphi[Q][gn][3]
 *            (used with per-point coefficients or isoparametric geometry)
 *   rtab     FP64 reduced tables wg[Qg] | dphi[Qg][n][3] | wm[Qm] | phi[Qm][n]
 *            for affine elements with element-constant coefficients (any
 *            rule whose sums sum_q wg dphi_a dphi_b^T and sum_q wm phi_a phi_b
 *            equal the basis rule's); rtab32 the same in FP32 (precision 1)
 *   c_*      per element (E) [tensor (E,3,3)] when per_qp == 0, or per
 *            quadrature point (E,Q) [(E,Q,3,3)] when per_qp == 1; NULL = absent
 *   precision 0: FP64 throughout; 1: element products in FP32, geometry and
 *            CSR reduction in FP64 (element-constant coefficients only).
 *   flags    SKFA_FLAG_SYMMETRIC_TENSOR: the caller guarantees c_tensor is
 *            symmetric per element (a smaller per-element payload is used).
 * The element-constant path runs a per-element payload pass into a
 * plan-owned buffer (allocated on the first such call) and the row-gather
 * kernel; the other path runs one fused kernel.
 */
int skfa_assemble_bilinear(skfa_plan* plan, const double* xyz, int64_t npoints, const int32_t* vconn,
                           const int32_t* gconn, int32_t gnodes, const double* tables, int32_t nq,
                           const double* rtab, const float* rtab32, int32_t nqg, int32_t nqm,
                           const double* c_value, const double* c_gradient, const double* c_symgrad,
                           const double* c_divergence, const double* c_tensor, int32_t per_qp, int32_t precision,
                           int32_t flags, double* values, void* stream);
#define SKFA_FLAG_SYMMETRIC_TENSOR 1
/* SKFA_FLAG_VECTOR3_DIAGONAL: `plan` is the scalar (C = 1) plan of a
 * 3-component field's node connectivity and the form has only value /
 * gradient coefficients (element-constant, affine elements).  Its element
 * matrix is [c == d] times the scalar one (bilinear_form_assembler.cpp:
 * 83-89 with no symmetric-gradient or divergence term), so the scalar gather
 * runs and writes each entry as the 3x3 diagonal node block of the
 * 3-component DOF-level CSR (the C = 3 plan's pattern).  `values` then holds
 * 9 x the scalar plan's nnz. */
#define SKFA_FLAG_VECTOR3_DIAGONAL 2
/* SKFA_FLAG_BARYCENTRIC_TABLE: for 10-node elements, rtab (and rtab32)
 * continue after pm with c[16][Qg]: the barycentric coefficients of the
 * reference gradients at the Qg gradient points, grad phi_b = sum_m c_bm
 * grad L_m, pairs in the order (0,0) (1,1) (2,2) (3,3), then for each edge
 * e = (i,j) of ((0,1),(1,2),(2,0),(0,3),(1,3),(2,3)) the pairs (4+e, i) and
 * (4+e, j).  Then, from the next multiple of 4 entries, the pair table
 * pt[10][10][8]: pt[a][b][3i:3i+3] = q_(ab,i) = sum_q wg_q c_(a,i)(q)
 * grad phi_b(q) (i = 0, 1; the c row of a's i-th pair, zero for i = 1 at a
 * vertex), pt[a][b][6] = sum_q wm_q phi_a phi_b, pt[a][b][7] = 0; so that
 * sum_q wg_q grad phi_a grad phi_b^T = sum_i grad L_m(a,i) q_(ab,i)^T.  The
 * scalar gather then forms element rows from the coefficient structure and
 * the vector-P2 gather its 3x3 node blocks from the pair table (tuning key
 * SKFA_TUNE_ROW_FORM). */
#define SKFA_FLAG_BARYCENTRIC_TABLE 4

/* skfa_assemble_bilinear followed by the load vector of a constant source
 * f (one device double) on the same scalar (C = 1) plan: `load` (nnodes
 * FP64, overwritten) receives what skfa_assemble_linear(plan, xyz, vconn,
 * gconn, gnodes, tables, nq, f, 0, NULL, 0, load) would, to rounding (the
 * row sums run in the row's incidence order).  Replaces the
 * BilinearFormAssembler::assemble + LinearFormAssembler::assemble pair
 * (cpp/src/assembly/bilinear_form_assembler.cpp:42-105,
 * cpp/src/assembly/linear_form_assembler.cpp:41-137) that the reference's
 * asm(bilinear) + asm(linear) calls run one after the other.  When the
 * element-table gather of 4-node elements runs, the gather lane that owns
 * node row i also sums |det J_e| w_a over the row's incidences
 * (w_a = sum_q W_q phi_a(X_q)), so the load costs no pass of its own;
 * otherwise the linear-form kernels run after the gather.  Vector fields
 * (components 3, or SKFA_FLAG_VECTOR3_DIAGONAL) are rejected (-1).
 * value_map / row_map (both or neither, int32 device arrays of the plan's
 * nnz and nnodes): entry k of the plan's CSR is written to
 * values[value_map[k]] and the load of node row i to load[row_map[i]],
 * entries with a negative map value are dropped, e.g. the Dirichlet
 * condensation K[I][:, I], b[I] of
 * examples/scikit-fem-gallery/comparison.py:80-87 straight from the gather
 * (the maps must be injective on the kept entries).  Without the fused
 * path the full results go through plan-owned buffers (allocated on first
 * use) and a scatter. */
int skfa_assemble_bilinear_load(skfa_plan* plan, const double* xyz, int64_t npoints, const int32_t* vconn,
                                const int32_t* gconn, int32_t gnodes, const double* tables, int32_t nq,
                                const double* rtab, const float* rtab32, int32_t nqg, int32_t nqm,
                                const double* c_value, const double* c_gradient, const double* c_symgrad,
                                const double* c_divergence, const double* c_tensor, int32_t per_qp,
                                int32_t precision, int32_t flags, double* values, const double* f, double* load,
                                const int32_t* value_map, const int32_t* row_map, void* stream);

/* skfa_assemble_bilinear without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as skfa_assemble_bilinear; the
 * per-element payload and the element tables of the element-constant path
 * live in a 96 MiB device array of the library (one per GPU, held until the
 * process ends) instead of plan-owned buffers.  A plan whose payload does not
 * fit (more than about 1.5 M elements with a mass term, 2 M without, 0.9 M
 * for 3-component fields) is rejected with an error before any launch.
 * Every call on one GPU uses the same array: calls that overlap on that GPU
 * (other streams, other threads) race. */
int skfa_assemble_bilinear_noalloc(skfa_plan* plan, const double* xyz, int64_t npoints, const int32_t* vconn,
                                   const int32_t* gconn, int32_t gnodes, const double* tables, int32_t nq,
                                   const double* rtab, const float* rtab32, int32_t nqg, int32_t nqm,
                                   const double* c_value, const double* c_gradient, const double* c_symgrad,
                                   const double* c_divergence, const double* c_tensor, int32_t per_qp,
                                   int32_t precision, int32_t flags, double* values, void* stream);

/* Linear form f.v + G:grad(v) into out (nrows FP64, overwritten).
 * Replaces LinearFormAssembler::assemble
 * (cpp/src/assembly/linear_form_assembler.cpp:41-137).  f: (C) constant
 * (f_per_qp 0) or (E,Q,C); G: (C,3) constant or (E,Q,C,3).  A constant f
 * without G on affine elements uses a per-element |det J| buffer that the
 * plan allocates on first use. */
int skfa_assemble_linear(skfa_plan* plan, const double* xyz, const int32_t* vconn, const int32_t* gconn,
                         int32_t gnodes, const double* tables, int32_t nq, const double* f, int32_t f_per_qp,
                         const double* G, int32_t G_per_qp, double* out, void* stream);

/* Facet (boundary) load vector into out (nnodes * components FP64,
 * overwritten).  Replaces LinearFormAssembler::assemble on the facet DOFs of
 * a FacetBasis (cpp/src/assembly/linear_form_assembler.cpp:41-137 driven by
 * python/skfemntv/linear_form.py:17-107 over basis.py:2442-2685).
 *   plan        skfa_plan_create on the (F, n) node connectivity of the
 *               facets' parent elements, one component (its incidence lists
 *               are used)
 *   face        (F) int32 row of face_shape for each facet (its corner
 *               orientation in the parent; 0 <= face < K)
 *   face_shape  (K, Q, n) FP64 parent shape functions at the facet points,
 *               one row per corner orientation (local face and the order
 *               of its corners, K <= 24)
 *   weights     (F, Q, components) FP64 dx * f
 * out[node * components + c] sums its (facet, local node) incidences in facet
 * order. */
int skfa_assemble_facet_linear(const skfa_plan* plan, const int32_t* face, const double* face_shape, int32_t nq,
                               const double* weights, double* out, int32_t components, void* stream);

/* Facet mass-type bilinear form (value kind: sum_q w phi_a phi_b on the
 * diagonal component blocks) into the CSR values of `plan` (nnz FP64,
 * overwritten).  Replaces BilinearFormAssembler::assemble with a value
 * coefficient on the facet DOFs of a FacetBasis
 * (cpp/src/assembly/bilinear_form_assembler.cpp:39-105 driven by
 * python/skfemntv/bilinear_form.py:13-99).
 *   plan        skfa_plan_create on the (F, n) node connectivity of the
 *               facets' parent elements, with the field's components (its
 *               pattern is the reference's facet pattern)
 *   conn        that (F, n) int32 connectivity
 *   face, face_shape as for skfa_assemble_facet_linear; weights (F, Q) dx * coefficient
 * Each node row adds its (facet, local node) incidences in facet order. */
int skfa_assemble_facet_bilinear(const skfa_plan* plan, const int32_t* conn, const int32_t* face,
                                 const double* face_shape, int32_t nq, const double* weights, double* values,
                                 void* stream);

/* ---------------------------------------------------------------------------
 * Cut-cell (ragged quadrature) forms.  Cell e owns the points
 * [offsets[e], offsets[e+1]) of flat per-point tables: weights (P), shape
 * values (P, n) and physical gradients (P, n, dim).  The plan is
 * skfa_cross_plan_create on the (cells, n) node connectivity of the cells
 * that have points (DOF = node * components + component), row and column
 * connectivity equal for the square forms; offsets are those cells' point
 * ranges (cells + 1 int64).
 *
 * Bilinear forms into the CSR values of the plan (nnz FP64, overwritten).
 * Replaces CutBilinearFormAssembler::assemble
 * (cpp/src/assembly/cut_form_assembler.cpp:136-205) and
 * CutCrossAssembler::assemble (cpp/src/assembly/cut_cross_assembler.cpp:70-137,
 * contraction cpp/include/native_fem/cross_contraction.hpp:42-97).
 *   terms   SKFA_CUT_VALUE and/or SKFA_CUT_GRADIENT: entry (a,c),(b,c) =
 *           w_q (value_q phi_a psi_b + gradient_q grad phi_a . grad psi_b)
 *           summed over the cell's points, other component pairs zero (value
 *           or gradient NULL = 1); needs equal component counts;
 *           SKFA_CUT_TENSOR: entry (a,r),(b,c) = w_q sum_ij h_ai T[q,r,(i),c,(j)]
 *           h_bj with h = phi or grad phi by row_kind / col_kind and T the
 *           (P, Cr, [row_dim], Cc, [col_dim]) tensor.
 *   one_lane_per_row  1 when a cell repeats a column node (one lane then adds
 *           all of a row's entries), else 0.
 * Each entry adds its cells in cell order; results are deterministic.  Value /
 * gradient terms with one_lane_per_row = 0 form per-cell node matrices
 * (cells x n_row x n_col FP64) in a buffer the plan keeps, allocated on the
 * first such call: calls on one plan must not overlap. */
#define SKFA_CUT_VALUE 1
#define SKFA_CUT_GRADIENT 2
#define SKFA_CUT_TENSOR 4
int skfa_assemble_cut_bilinear(const skfa_cross_plan* plan, const int64_t* offsets, const double* weights,
                               const double* row_shape, const double* row_grad, int32_t row_dim,
                               const double* col_shape, const double* col_grad, int32_t col_dim, int32_t terms,
                               const double* value, const double* gradient, int32_t row_kind, int32_t col_kind,
                               const double* tensor, int32_t one_lane_per_row, double* values, void* stream);

/* Cut-cell load vector out (row nodes * components FP64, overwritten):
 * out[node*C + c] = sum over the node's cells of sum_q w_q (phi_a(q) f_c(q) +
 * grad phi_a(q) . G_c(q)).  Replaces CutLinearFormAssembler::assemble
 * (cpp/src/assembly/cut_form_assembler.cpp:67-134).
 *   value     (C) constant or (P, C) per point (value_per_point = 1); NULL = no term
 *   gradient  (C, dim) constant or (P, C, dim) per point; NULL = no term
 * Per-cell node vectors (cells x n x C FP64) go through the plan's cut-cell
 * buffer (allocated or grown on first use): calls on one plan must not
 * overlap. */
int skfa_assemble_cut_linear(const skfa_cross_plan* plan, const int64_t* offsets, const double* weights,
                             const double* shape, const double* grad, int32_t dim, const double* value,
                             int32_t value_per_point, const double* gradient, int32_t gradient_per_point,
                             double* out, void* stream);

/* Per (e,q) geometry: replaces tabulate_basis_geometry
 * (cpp/src/assembly/basis_geometry.cpp:49-176).  Outputs (E,Q[,...]) FP64,
 * each optional (NULL = not written): det, tol (64 eps max|J|^3), cond,
 * dx = |det| w, x (E,Q,3), shape (E,Q,n), grads (E,Q,n,3).  Validation
 * (basis_geometry.cpp:178-205) is done by the caller from det/tol. */
int skfa_tabulate_geometry(const double* xyz, const int32_t* vconn, const int32_t* gconn, int32_t gnodes, int64_t E,
                           const double* tables, int32_t nq, int32_t n, double* det, double* tol, double* cond,
                           double* dx, double* x, double* shape, double* grads, void* stream);

/* ---------------------------------------------------------------------------
 * First-seen oriented edge numbering of tetrahedra.  Replaces
 * build_oriented_edge_topology (cpp/src/assembly/edge_topology.cpp:30-84) and
 * gives the edge-node order of MeshTet2.from_mesh
 * (python/skfemntv/basis.py:644-658).
 *   vconn (E,4) int32; elem_edges (E,6) int32 and signs (E,6) int8 out;
 *   edge_lo / edge_hi capacity 6E int64 out (may be NULL); *nedges (host) out;
 *   *bad_cell (host) = first cell repeating a corner vertex, else -1.
 */
int skfa_edge_topology(const int32_t* vconn, int64_t E, int64_t nvertices, int32_t* elem_edges, int8_t* signs,
                       int64_t* edge_lo, int64_t* edge_hi, int64_t* nedges, int64_t* bad_cell, void* stream);

/* Nedelec N1 mass + curl-curl on affine tetrahedra into CSR values of an edge
 * plan (n = 6, C = 1).  Replaces integrate_tet_n1 + the fixed-pattern
 * np.add.at scatter (cpp/src/assembly/edge_topology.cpp:156-233,
 * python/skfemntv/hcurl_assembler.py:94-97,181-212).
 *   tables FP64 w[Q] | phi[Q][6][3] (reference Whitney values) | curl[6][3];
 *   kappa / mu per element (per_qp 0) or per (E,Q). */
int skfa_assemble_tet_n1(skfa_plan* plan, const double* xyz, const int32_t* vconn, const double* tables,
                         int32_t nq, const double* kappa, const double* mu, int32_t per_qp, int32_t include_mass,
                         int32_t include_curl, double* values, void* stream);

/* HIP event helpers for timing on the launch stream. */
int skfa_event_create(void** ev);
int skfa_event_destroy(void* ev);
int skfa_event_record(void* ev, void* stream);
int skfa_event_elapsed_ms(void* start, void* stop, float* ms);
int skfa_stream_synchronize(void* stream);

/* ---------------------------------------------------------------------------
 * Solving assembled systems (batched, block-diagonal CSR).
 *
 * No native counterpart in the reference: its solver policy is external
 * (python/skfemntv/capabilities.py:67).  The Dirichlet condensation mirrors
 * the reference's only solve pattern, solve_dirichlet in
 * examples/scikit-fem-gallery/comparison.py:80-87
 * (free = setdiff1d(arange(n), boundary); matrix[free][:, free]).
 * A batch of independent systems is one CSR whose system s owns rows
 * [sys_ptr[s], sys_ptr[s+1]) and has no couplings outside them.
 * ------------------------------------------------------------------------- */

/* y = A x (FP64), `lanes_per_row` in {4, 8, 16}. */
int skfa_spmv(const int64_t* indptr, const int32_t* indices, const double* vals, const double* x, double* y,
              int64_t nrows, int32_t lanes_per_row, void* stream);

/* Keep rows/columns with keep[i] != 0, in order.  newid (nrows+1) receives
 * the exclusive scan of keep (new index of kept row i); out_* are sized by
 * the caller for the full input (nnz and nrows+1 suffice); *out_nrows and
 * *out_nnz (host) receive the sizes.  Synchronises `stream`. */
int skfa_csr_condense(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t nrows,
                      const uint8_t* keep, int64_t* newid, int64_t* out_indptr, int32_t* out_indices,
                      double* out_vals, int64_t* out_nrows, int64_t* out_nnz, void* stream);

/* Lower-triangular (j <= i) pattern of a CSR: lptr (nrows+1), and when lidx
 * and lsrc are non-null the column indices and source positions.  *nnz_lower
 * is a host output.  Call once with lidx = lsrc = NULL to size. */
int skfa_csr_lower(const int64_t* indptr, const int32_t* indices, int64_t nrows, int64_t* lptr, int32_t* lidx,
                   int64_t* lsrc, int64_t* nnz_lower, void* stream);

/* Transpose pattern: tptr (ncols+1), tidx (nnz) sorted per row, tsrc (nnz)
 * = position of each entry in the input.  Deterministic. */
int skfa_csr_transpose(const int64_t* ptr, const int32_t* idx, int64_t nrows, int64_t ncols, int64_t* tptr,
                       int32_t* tidx, int64_t* tsrc, void* stream);

/* Transpose of the lower pattern (lptr, lidx) of a CSR whose pattern is
 * symmetric (FEM matrices): tptr (nrows+1), tidx, tsrc as skfa_csr_transpose,
 * without atomics or sorting.  Returns -1 if the pattern is not symmetric.
 * Synchronises `stream`. */
int skfa_csr_transpose_lower(const int64_t* indptr, const int32_t* indices, int64_t nrows, const int64_t* lptr,
                             const int32_t* lidx, int64_t* tptr, int32_t* tidx, int64_t* tsrc, void* stream);

/* out[k] = src[map[k]] (FP64 and FP32). */
int skfa_gather_values(const double* src, const int64_t* map, int64_t n, double* out, void* stream);
int skfa_gather_values_f32(const float* src, const int64_t* map, int64_t n, float* out, void* stream);

/* dinv[i] = 1 / A_ii. */
int skfa_diag_inverse(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t n, double* dinv,
                      void* stream);

/* Per-system sums: out_xy[s] = sum x_i y_i, out_xx[s] = sum x_i^2 (optional). */
int skfa_sys_dot(const double* x, const double* y, const int64_t* sys_ptr, int32_t nsys, double* out_xy,
                 double* out_xx, void* stream);

/* Batched preconditioned conjugate gradients, FP64, x0 = 0.  kind 0: Jacobi
 * (dinv); kind 1: z = G^T (G r) with G lower CSR (FP32 values) and its
 * transpose.  System s stops when ||r_s|| <= tol ||b_s|| (recursive
 * residual) or after max_iter; converged systems stop costing bandwidth
 * (their row blocks return at once).  Workspaces (device): work 7n doubles,
 * sys_work 6*nsys doubles, int_work nsys+1 int32.  iters (device, nsys)
 * receives per-system iteration counts, *total_iters (host) the number of
 * batch iterations run.  Reductions have a fixed order (reproducible). */
int skfa_pcg_solve(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t n,
                   const int64_t* sys_ptr, int32_t nsys, const double* b, double* x, int32_t kind, const double* dinv,
                   const int64_t* gptr, const int32_t* gidx, const float* gval, const int64_t* gtptr,
                   const int32_t* gtidx, const float* gtval, double tol, int32_t max_iter, int32_t check_every,
                   int32_t spmv_lanes, double* work, double* sys_work, int32_t* int_work, int32_t* iters,
                   int32_t* total_iters, void* stream);

/* ---------------------------------------------------------------------------
 * Learned FSAI factor (no counterpart in the reference).
 *
 * At = D^{-1/2} A D^{-1/2}; the factor G = Gt D^{-1/2} has A's lower pattern;
 * a 16-128-128-1 ReLU MLP maps 16 features of each lower nonzero to an output
 * o (Gt_ij = o_ij off the diagonal, Gt_ii = exp(o_ii)).  Parameters are one
 * FP32 vector of SKFA_FSAI_PARAMS floats: W1[16][128] | b1[128] |
 * W2[128][128] | b2[128] | w3[128] | b3.  MFMA operands are BF16 copies
 * ("fragments", SKFA_FSAI_FRAG uint16) made by skfa_mlp_pack / skfa_mlp_adam.
 * ------------------------------------------------------------------------- */
#define SKFA_FSAI_FEATURES 16
#define SKFA_FSAI_PARAMS 18817
#define SKFA_FSAI_FRAG 34816

/* Features (float, nnz_lower x 16) of every lower nonzero; also writes
 * rsqrt_diag (n) = A_ii^{-1/2} and row_stats (4n doubles).  xyz (n x 3
 * doubles, coordinates of each unknown) may be NULL (geometric features 0). */
int skfa_fsai_features(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t n,
                       const double* xyz, const int64_t* lptr, const int32_t* lidx, double* rsqrt_diag,
                       double* row_stats, float* features, void* stream);

/* Factor values from network outputs: gtilde (FP64, optional), g = Gt D^{-1/2}
 * (FP32, optional), logdiag (n, optional). */
int skfa_fsai_factor(const float* outputs, const int64_t* lptr, const int32_t* lidx, const double* rsqrt_diag,
                     int64_t n, double* gtilde, float* g, double* logdiag, void* stream);

/* Kaporin objective per system, loss[s] = log(tr(Gt At Gt^T)/n_s)
 * - (2/n_s) sum log Gt_ii, and (if grad_outputs != NULL) grad_scale * dL/do
 * per lower nonzero.  work: 3*nnz_lower + n + 3*nsys + 1 doubles; row_sys: n
 * int32.  Synchronises `stream` once (reads nnz_lower). */
int skfa_fsai_kaporin(const int64_t* indptr, const int32_t* indices, const double* vals, int64_t n,
                      const int64_t* sys_ptr, int32_t nsys, const int64_t* lptr, const int32_t* lidx,
                      const double* rsqrt_diag, const float* outputs, double grad_scale, double* work,
                      int32_t* row_sys, double* loss, float* grad_outputs, void* stream);

/* BF16 MFMA fragments from FP32 parameters. */
int skfa_mlp_pack(const float* params, uint16_t* frag, void* stream);

/* out[k] = MLP(features[k]) for k < nnz (BF16 MFMA, FP32 accumulate). */
int skfa_mlp_forward(const float* params, const uint16_t* frag, const float* features, int64_t nnz, float* out,
                     void* stream);

/* Workspace bytes skfa_mlp_backward needs for nnz rows. */
size_t skfa_mlp_workspace_bytes(int64_t nnz);

/* grad (SKFA_FSAI_PARAMS floats) = sum_k grad_out[k] * d MLP(features[k]) / d params.
 * Deterministic (fixed partial-sum order). */
int skfa_mlp_backward(const float* params, const uint16_t* frag, const float* features, const float* grad_out,
                      int64_t nnz, void* workspace, size_t workspace_bytes, float* grad, void* stream);

/* One Adam step on the FP32 parameters (gradient multiplied by clip_scale),
 * then repack the BF16 fragments. */
int skfa_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                  float eps, int32_t step, float clip_scale, uint16_t* frag, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SKFEM_AMD_H */
