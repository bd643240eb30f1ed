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
int skfa_plan_destroy(skfa_plan* plan);
/* Launch tuning of an existing plan (results are identical for every
 * setting; only speed changes).  Keys: */
#define SKFA_TUNE_XCD_CHUNK 1  /* waves per XCD chunk, 0 = plain dispatch order */
#define SKFA_TUNE_LDS_ATOMIC 2 /* 1: accumulate with ds_add_f64, 0: read-add-write */
#define SKFA_TUNE_PLACEMENT 3  /* 0 plain order, 1 one contiguous segment per XCD, 2 XCD chunks */
#define SKFA_TUNE_STREAMING 4  /* 1: non-temporal record loads and CSR value stores */
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

/* Linear form f.v + G:grad(v) into out (nrows FP64, overwritten).
 * Replaces LinearFormAssembler::assemble
 * (cpp/src/assembly/linear_form_assembler.cpp:41-137).  f: (C) constant
 * (f_per_qp 0) or (E,Q,C); G: (C,3) constant or (E,Q,C,3). */
int skfa_assemble_linear(const skfa_plan* plan, const double* xyz, const int32_t* vconn, const int32_t* gconn,
                         int32_t gnodes, const double* tables, int32_t nq, const double* f, int32_t f_per_qp,
                         const double* G, int32_t G_per_qp, double* out, void* stream);

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

#ifdef __cplusplus
}
#endif

#endif /* SKFEM_AMD_H */
