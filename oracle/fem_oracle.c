/*
 * CPU oracle for the element-assembly path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the scikit-fem-native (skfemntv 0.3.1) CPU
 * algorithms that the MI355X library replaces.  Only tests/, the smoke
 * check and the bench's CPU-baseline leg load this file's library; the
 * product never links it.  Every function cites the reference code it
 * restates (paths relative to the reference repository root).
 *
 * Arrays are C-contiguous; int64 indices; FP64 values.  Functions return 0
 * on success and a negative code on invalid input; oracle_last_error()
 * gives the message.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

static char g_error[512];
const char* oracle_last_error(void) { return g_error; }
static int fail(const char* msg) { snprintf(g_error, sizeof g_error, "%s", msg); return -1; }
void oracle_free(void* p) { free(p); }

/* ------------------------------------------------------------------ */
/* Geometry: cpp/src/assembly/basis_geometry.cpp:18-47 (det/adjugate)   */
/* ------------------------------------------------------------------ */
static double det3(const double* m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}
static double det2(const double* m) { return m[0] * m[3] - m[1] * m[2]; }

static void inv_by_adjugate(const double* m, double* r, int d) {
    if (d == 2) {
        const double det = det2(m);
        r[0] = m[3] / det; r[1] = -m[1] / det; r[2] = -m[2] / det; r[3] = m[0] / det;
        return;
    }
    const double det = det3(m);
    r[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    r[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    r[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    r[3] = (m[5] * m[6] - m[3] * m[8]) / det;
    r[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    r[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    r[6] = (m[3] * m[7] - m[4] * m[6]) / det;
    r[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    r[8] = (m[0] * m[4] - m[1] * m[3]) / det;
}

/*
 * tabulate_basis_geometry, cpp/src/assembly/basis_geometry.cpp:49-209.
 * p (d,P) component-major, t (gn,E) node-major, shape (Q,n), refgrad (Q,n,d),
 * gshape (Q,gn), ggrad (Q,gn,d), W (Q).  Outputs (E,Q,...) as in the
 * reference.  The validation pass (:178-205) reports the first failing
 * (cell, qp) through bad_cell, bad_qp, bad_reason (0 none, 1 near-singular,
 * 2 orientation change) instead of raising.
 */
int oracle_tabulate_geometry(const double* p, int64_t P, int d, const int64_t* t, int64_t gn, int64_t E,
                             const double* shape, const double* refgrad, int64_t Q, int64_t n,
                             const double* gshape, const double* ggrad, const double* W,
                             double* out_shape, double* out_grad, double* out_dx, double* out_x,
                             double* out_det, double* out_tol, double* out_cond,
                             int64_t* bad_cell, int64_t* bad_qp, int* bad_reason) {
    if (d != 2 && d != 3) return fail("dimension must be 2 or 3");
    *bad_cell = -1; *bad_qp = -1; *bad_reason = 0;
    for (int64_t e = 0; e < E; ++e) {
        for (int64_t q = 0; q < Q; ++q) {
            double J[9] = {0}, Ji[9] = {0}, x[3] = {0};
            for (int64_t a = 0; a < gn; ++a) {
                const int64_t v = t[a * E + e];
                if (v < 0 || v >= P) return fail("connectivity index is out of bounds");
                for (int i = 0; i < d; ++i) {
                    const double c = p[i * P + v];
                    x[i] += gshape[q * gn + a] * c;
                    for (int j = 0; j < d; ++j) J[i * d + j] += c * ggrad[(q * gn + a) * d + j];
                }
            }
            for (int i = 0; i < d; ++i) out_x[(e * Q + q) * d + i] = x[i];
            double scale = 0.;
            for (int i = 0; i < d * d; ++i) scale = fmax(scale, fabs(J[i]));
            const double sp = d == 2 ? scale * scale : scale * scale * scale;
            const double tol = 64. * DBL_EPSILON * sp;
            const double det = d == 2 ? det2(J) : det3(J);
            out_det[e * Q + q] = det;
            out_tol[e * Q + q] = tol;
            out_dx[e * Q + q] = 0.; out_cond[e * Q + q] = 0.;
            if (!isfinite(det) || !isfinite(tol) || !(fabs(det) > tol)) continue;
            inv_by_adjugate(J, Ji, d);
            double jn = 0., in = 0.;
            for (int i = 0; i < d * d; ++i) { jn += J[i] * J[i]; in += Ji[i] * Ji[i]; }
            out_cond[e * Q + q] = sqrt(jn * in);
            out_dx[e * Q + q] = fabs(det) * W[q];
            for (int64_t a = 0; a < n; ++a) {
                out_shape[(e * Q + q) * n + a] = shape[q * n + a];
                for (int j = 0; j < d; ++j) {
                    double s = 0.;
                    for (int k = 0; k < d; ++k) s += refgrad[(q * n + a) * d + k] * Ji[k * d + j];
                    out_grad[((e * Q + q) * n + a) * d + j] = s;
                }
            }
        }
    }
    for (int64_t e = 0; e < E && *bad_reason == 0; ++e) {
        for (int64_t q = 0; q < Q; ++q) {
            const double det = out_det[e * Q + q], tol = out_tol[e * Q + q];
            if (isfinite(det) && isfinite(tol) && fabs(det) > tol) continue;
            *bad_cell = e; *bad_qp = q; *bad_reason = 1;
            break;
        }
        if (*bad_reason) break;
        const int neg = out_det[e * Q] < 0.;
        for (int64_t q = 1; q < Q; ++q) {
            if ((out_det[e * Q + q] < 0.) == neg) continue;
            *bad_cell = e; *bad_qp = q; *bad_reason = 2;
            break;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* CSR pattern + element->CSR map                                      */
/* cpp/src/assembly/bilinear_form_assembler.cpp:111-120 (square),      */
/* cross_bilinear_assembler.cpp:307-320 (rectangular),                 */
/* edge_topology.cpp:86-154 (edge pattern, same rule).                 */
/* ------------------------------------------------------------------ */
static int cmp_i64(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* rows (E,Lr) and cols (E,Lc) entity-major.  Returns malloc'ed arrays. */
int oracle_pattern_build(const int64_t* rows, const int64_t* cols, int64_t E, int64_t Lr, int64_t Lc,
                         int64_t nrows, int64_t** indptr_out, int64_t** indices_out,
                         int64_t** scatter_out, int64_t* nnz_out) {
    int64_t* count = calloc((size_t)nrows + 1, sizeof(int64_t));
    if (!count) return fail("out of memory");
    for (int64_t e = 0; e < E; ++e)
        for (int64_t i = 0; i < Lr; ++i) {
            const int64_t r = rows[e * Lr + i];
            if (r < 0 || r >= nrows) { free(count); return fail("row index out of bounds"); }
            count[r + 1] += Lc;
        }
    for (int64_t r = 0; r < nrows; ++r) count[r + 1] += count[r];
    int64_t* cand = malloc(sizeof(int64_t) * (size_t)(count[nrows] ? count[nrows] : 1));
    int64_t* fillp = malloc(sizeof(int64_t) * (size_t)(nrows ? nrows : 1));
    if (!cand || !fillp) { free(count); free(cand); free(fillp); return fail("out of memory"); }
    memcpy(fillp, count, sizeof(int64_t) * (size_t)nrows);
    for (int64_t e = 0; e < E; ++e)
        for (int64_t i = 0; i < Lr; ++i) {
            const int64_t r = rows[e * Lr + i];
            for (int64_t j = 0; j < Lc; ++j) cand[fillp[r]++] = cols[e * Lc + j];
        }
    int64_t* indptr = malloc(sizeof(int64_t) * (size_t)(nrows + 1));
    indptr[0] = 0;
    int64_t nnz = 0;
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t* seg = cand + count[r];
        const int64_t m = count[r + 1] - count[r];
        qsort(seg, (size_t)m, sizeof(int64_t), cmp_i64);
        int64_t u = 0;
        for (int64_t k = 0; k < m; ++k)
            if (k == 0 || seg[k] != seg[k - 1]) cand[nnz + u++] = seg[k];
        nnz += u;
        indptr[r + 1] = nnz;
    }
    int64_t* indices = malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
    memcpy(indices, cand, sizeof(int64_t) * (size_t)nnz);
    free(cand); free(fillp); free(count);
    int64_t* scatter = malloc(sizeof(int64_t) * (size_t)((E * Lr * Lc) > 0 ? E * Lr * Lc : 1));
    for (int64_t e = 0; e < E; ++e)
        for (int64_t i = 0; i < Lr; ++i) {
            const int64_t r = rows[e * Lr + i];
            const int64_t lo0 = indptr[r], hi0 = indptr[r + 1];
            for (int64_t j = 0; j < Lc; ++j) {
                const int64_t c = cols[e * Lc + j];
                int64_t lo = lo0, hi = hi0; /* lower_bound */
                while (lo < hi) {
                    const int64_t mid = lo + (hi - lo) / 2;
                    if (indices[mid] < c) lo = mid + 1; else hi = mid;
                }
                scatter[(e * Lr + i) * Lc + j] = lo;
            }
        }
    *indptr_out = indptr; *indices_out = indices; *scatter_out = scatter; *nnz_out = nnz;
    return 0;
}

/* Greedy first-fit element colouring over row DOFs,                  */
/* bilinear_form_assembler.cpp:121-140 / cross_bilinear_assembler.cpp:321-341.
 * colour_out (E) receives each element's colour; returns colour count.  */
int64_t oracle_colouring(const int64_t* rows, int64_t E, int64_t Lr, int64_t nrows, int32_t* colour_out) {
    /* per-row list of colours already used, stored as a bitset per row of up to 1024 colours */
    const int64_t words = 16; /* 1024 colours */
    uint64_t* used = calloc((size_t)(nrows * words), sizeof(uint64_t));
    int64_t ncol = 0;
    for (int64_t e = 0; e < E; ++e) {
        uint64_t mark[16] = {0};
        for (int64_t i = 0; i < Lr; ++i) {
            const int64_t r = rows[e * Lr + i];
            for (int64_t w = 0; w < words; ++w) mark[w] |= used[r * words + w];
        }
        int64_t c = 0;
        while (c < 1024 && (mark[c >> 6] >> (c & 63)) & 1ull) ++c;
        colour_out[e] = (int32_t)c;
        if (c + 1 > ncol) ncol = c + 1;
        for (int64_t i = 0; i < Lr; ++i) {
            const int64_t r = rows[e * Lr + i];
            used[r * words + (c >> 6)] |= 1ull << (c & 63);
        }
    }
    free(used);
    return ncol;
}

/* ------------------------------------------------------------------ */
/* Scalar-kind bilinear element kernel,                                */
/* bilinear_form_assembler.cpp:79-105 (value/gradient/sym-grad/div).   */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *shape, *grads, *dx;           /* (E,Q,n) (E,Q,n,d) (E,Q) */
    const double *value, *gradient, *symgrad, *divergence; /* (E,Q) or NULL */
    const int64_t* scatter;                     /* (E,L,L) */
    double* values;
    int64_t n, C, Q, d;
} bilin_args;

static void bilin_element(const bilin_args* A, int64_t e) {
    const int64_t n = A->n, C = A->C, Q = A->Q, d = A->d, L = n * C;
    for (int64_t q = 0; q < Q; ++q) {
        const int64_t eq = e * Q + q;
        const double* g = A->grads + eq * n * d;
        const double* s = A->shape + eq * n;
        for (int64_t a = 0; a < n; ++a)
            for (int64_t b = 0; b < n; ++b) {
                double diag = 0.;
                if (A->value) diag += A->value[eq] * s[a] * s[b];
                if (A->gradient)
                    for (int64_t j = 0; j < d; ++j) diag += A->gradient[eq] * g[a * d + j] * g[b * d + j];
                for (int64_t c = 0; c < C; ++c)
                    for (int64_t dd = 0; dd < C; ++dd) {
                        double entry = c == dd ? diag : 0.;
                        if (A->symgrad) {
                            double con = g[a * d + dd] * g[b * d + c];
                            if (c == dd)
                                for (int64_t j = 0; j < d; ++j) con += g[a * d + j] * g[b * d + j];
                            entry += .5 * A->symgrad[eq] * con;
                        }
                        if (A->divergence) entry += A->divergence[eq] * g[a * d + c] * g[b * d + dd];
                        const int64_t i = a * C + c, k = b * C + dd;
                        A->values[A->scatter[(e * L + i) * L + k]] += entry * A->dx[eq];
                    }
            }
    }
}

typedef struct { const bilin_args* args; const int64_t* list; int64_t begin, end; int kind; const void* extra; } work_t;

/* ------------------------------------------------------------------ */
/* Tensor (anisotropic) cross kernel, cross_bilinear_assembler.cpp:    */
/* 207-247 + cross_contraction.hpp:42-84, gradient x gradient, scalar  */
/* fields: K[a,b] += dx * sum_ij grad_i(phi_a) A[e,q,i,j] grad_j(phi_b) */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *grads, *dx, *tensor; /* (E,Q,n,d) (E,Q) (E,Q,d,d) */
    const int64_t* scatter;
    double* values;
    int64_t n, Q, d;
} tensor_args;

static void tensor_element(const tensor_args* A, int64_t e) {
    const int64_t n = A->n, Q = A->Q, d = A->d;
    for (int64_t q = 0; q < Q; ++q) {
        const int64_t eq = e * Q + q;
        const double* g = A->grads + eq * n * d;
        const double* T = A->tensor + eq * d * d;
        for (int64_t a = 0; a < n; ++a)
            for (int64_t b = 0; b < n; ++b) {
                double r = 0.;
                for (int64_t i = 0; i < d; ++i)
                    for (int64_t j = 0; j < d; ++j) r += g[a * d + i] * T[i * d + j] * g[b * d + j];
                A->values[A->scatter[(e * n + a) * n + b]] += A->dx[eq] * r;
            }
    }
}

static void* worker(void* p) {
    work_t* w = (work_t*)p;
    for (int64_t k = w->begin; k < w->end; ++k) {
        const int64_t e = w->list ? w->list[k] : k;
        if (w->kind == 0) bilin_element(w->args, e);
        else tensor_element((const tensor_args*)w->extra, e);
    }
    return NULL;
}

/* Runs the element loop serially (threads<=1) or per colour with a static
 * contiguous split over `threads` pthreads, like parallel_for_workers
 * (cpp/include/native_fem/parallel.hpp:21-45). */
static void run_coloured(int kind, const bilin_args* b, const tensor_args* t, int64_t E,
                         const int32_t* colour, int64_t ncol, int threads) {
    if (threads <= 1 || !colour) {
        for (int64_t e = 0; e < E; ++e) {
            if (kind == 0) bilin_element(b, e); else tensor_element(t, e);
        }
        return;
    }
    int64_t* order = malloc(sizeof(int64_t) * (size_t)(E ? E : 1));
    int64_t* start = calloc((size_t)ncol + 1, sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e) start[colour[e] + 1]++;
    for (int64_t c = 0; c < ncol; ++c) start[c + 1] += start[c];
    int64_t* fillp = malloc(sizeof(int64_t) * (size_t)(ncol ? ncol : 1));
    memcpy(fillp, start, sizeof(int64_t) * (size_t)ncol);
    for (int64_t e = 0; e < E; ++e) order[fillp[colour[e]]++] = e;
    pthread_t th[256];
    work_t w[256];
    if (threads > 256) threads = 256;
    for (int64_t c = 0; c < ncol; ++c) {
        const int64_t cnt = start[c + 1] - start[c];
        int nt = threads;
        if (cnt < 128) nt = 1; /* serial below the explicit threshold, parallel.hpp:22-23 */
        for (int k = 0; k < nt; ++k) {
            w[k].args = b; w[k].extra = t; w[k].kind = kind; w[k].list = order + start[c];
            w[k].begin = cnt * k / nt; w[k].end = cnt * (k + 1) / nt;
            if (nt == 1) worker(&w[k]); else pthread_create(&th[k], NULL, worker, &w[k]);
        }
        if (nt > 1) for (int k = 0; k < nt; ++k) pthread_join(th[k], NULL);
    }
    free(order); free(start); free(fillp);
}

/* BilinearFormAssembler::assemble, bilinear_form_assembler.cpp:39-74.
 * values (nnz) is zeroed first.  colour may be NULL for the serial path. */
int oracle_assemble_bilinear(const double* shape, const double* grads, const double* dx,
                             const double* value, const double* gradient, const double* symgrad,
                             const double* divergence, const int64_t* scatter, int64_t E, int64_t n,
                             int64_t C, int64_t Q, int64_t d, double* values, int64_t nnz,
                             const int32_t* colour, int64_t ncol, int threads) {
    if (!value && !gradient && !symgrad && !divergence) return fail("at least one coefficient is required");
    memset(values, 0, sizeof(double) * (size_t)nnz);
    bilin_args a = {shape, grads, dx, value, gradient, symgrad, divergence, scatter, values, n, C, Q, d};
    run_coloured(0, &a, NULL, E, colour, ncol, threads);
    return 0;
}

/* CrossBilinearAssembler::assemble, gradient/gradient tensor path. */
int oracle_assemble_tensor(const double* grads, const double* dx, const double* tensor,
                           const int64_t* scatter, int64_t E, int64_t n, int64_t Q, int64_t d,
                           double* values, int64_t nnz, const int32_t* colour, int64_t ncol, int threads) {
    memset(values, 0, sizeof(double) * (size_t)nnz);
    tensor_args t = {grads, dx, tensor, scatter, values, n, Q, d};
    run_coloured(1, NULL, &t, E, colour, ncol, threads);
    return 0;
}

/* ------------------------------------------------------------------ */
/* General cross kernel between two fields, serial element loop:       */
/* CrossBilinearAssembler::assemble_entity (cross_bilinear_assembler.  */
/* cpp:207-247) with contract_cross_basis / contract_scalar_cross_basis */
/* (cross_contraction.hpp:42-97).  Row = test field (nr nodes, Cr      */
/* components), column = trial field (nc, Cc).  row_kind / col_kind    */
/* 0 value, 1 gradient.  coefficient: NULL (= 1) or scalar (E,Q) when  */
/* tensor == 0; (E,Q,Cr,[d],Cc,[d]) when tensor == 1.  scatter is the  */
/* rectangular map of oracle_pattern_build on (E,nr*Cr) x (E,nc*Cc).   */
/* ------------------------------------------------------------------ */
int oracle_assemble_cross(const double* rshape, const double* rgrads, const double* cshape,
                          const double* cgrads, const double* dx, const double* coefficient, int tensor,
                          int row_kind, int col_kind, const int64_t* scatter, int64_t E, int64_t nr, int64_t Cr,
                          int64_t nc, int64_t Cc, int64_t Q, int64_t d, double* values, int64_t nnz) {
    if (!tensor && Cr != Cc) return fail("scalar cross coefficient requires equal component counts");
    if (!tensor && row_kind != col_kind) return fail("value-gradient coupling requires a tensor coefficient");
    if (tensor && !coefficient) return fail("tensor coefficient is null");
    memset(values, 0, sizeof(double) * (size_t)nnz);
    const int64_t Lr = nr * Cr, Lc = nc * Cc;
    const int64_t Dr = row_kind ? d : 1, Dc = col_kind ? d : 1;
    for (int64_t e = 0; e < E; ++e)
        for (int64_t q = 0; q < Q; ++q) {
            const int64_t eq = e * Q + q;
            const double scale = dx[eq];
            for (int64_t a = 0; a < nr; ++a)
                for (int64_t b = 0; b < nc; ++b) {
                    if (!tensor) {
                        const double material = coefficient ? coefficient[eq] : 1.;
                        double contract = 0.;
                        if (!row_kind) {
                            contract = rshape[eq * nr + a] * cshape[eq * nc + b];
                        } else {
                            for (int64_t i = 0; i < d; ++i)
                                contract += rgrads[(eq * nr + a) * d + i] * cgrads[(eq * nc + b) * d + i];
                        }
                        const double entry = scale * material * contract;
                        for (int64_t r = 0; r < Cr; ++r)
                            values[scatter[(e * Lr + a * Cr + r) * Lc + b * Cc + r]] += entry;
                    } else {
                        for (int64_t r = 0; r < Cr; ++r)
                            for (int64_t c = 0; c < Cc; ++c) {
                                double res = 0.;
                                for (int64_t i = 0; i < Dr; ++i)
                                    for (int64_t j = 0; j < Dc; ++j) {
                                        const double rv = row_kind ? rgrads[(eq * nr + a) * d + i]
                                                                   : rshape[eq * nr + a];
                                        const double cv = col_kind ? cgrads[(eq * nc + b) * d + j]
                                                                   : cshape[eq * nc + b];
                                        const double t =
                                            coefficient[(((eq * Cr + r) * Dr + i) * Cc + c) * Dc + j];
                                        res += rv * t * cv;
                                    }
                                values[scatter[(e * Lr + a * Cr + r) * Lc + b * Cc + c]] += scale * res;
                            }
                    }
                }
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Linear form, linear_form_assembler.cpp:103-137 (serial path).       */
/* value: (C) constant or (E,Q,C); gradient: (C,d) or (E,Q,C,d).       */
/* ------------------------------------------------------------------ */
int oracle_assemble_linear(const double* shape, const double* grads, const double* dx,
                           const int64_t* dofs, int64_t E, int64_t n, int64_t C, int64_t Q, int64_t d,
                           const double* value, int const_value, const double* gradient, int const_gradient,
                           double* out, int64_t ndofs) {
    if (!value && !gradient) return fail("at least one coefficient is required");
    memset(out, 0, sizeof(double) * (size_t)ndofs);
    if ((!value || const_value) && (!gradient || const_gradient)) {
        for (int64_t e = 0; e < E; ++e)
            for (int64_t a = 0; a < n; ++a)
                for (int64_t c = 0; c < C; ++c) {
                    double local = 0.;
                    for (int64_t q = 0; q < Q; ++q) {
                        double con = 0.;
                        if (value) con += shape[(e * Q + q) * n + a] * value[c];
                        if (gradient)
                            for (int64_t j = 0; j < d; ++j)
                                con += grads[((e * Q + q) * n + a) * d + j] * gradient[c * d + j];
                        local += dx[e * Q + q] * con;
                    }
                    out[dofs[(e * n + a) * C + c]] += local;
                }
        return 0;
    }
    for (int64_t e = 0; e < E; ++e)
        for (int64_t q = 0; q < Q; ++q)
            for (int64_t a = 0; a < n; ++a)
                for (int64_t c = 0; c < C; ++c) {
                    double con = 0.;
                    const int64_t eq = e * Q + q;
                    if (value) con += shape[eq * n + a] * value[const_value ? c : eq * C + c];
                    if (gradient)
                        for (int64_t j = 0; j < d; ++j)
                            con += grads[(eq * n + a) * d + j] *
                                   gradient[const_gradient ? c * d + j : (eq * C + c) * d + j];
                    out[dofs[(e * n + a) * C + c]] += dx[eq] * con;
                }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Oriented edge topology, edge_topology.cpp:30-84 (tetrahedra).       */
/* First-seen ids over (cell, local edge); sign +1 iff the directed    */
/* local pair starts at the smaller global vertex.  Uses an open-      */
/* addressing hash table instead of unordered_map (same ids).          */
/* ------------------------------------------------------------------ */
int oracle_edge_topology(const int64_t* t, int64_t E, int64_t* edges_out /* (2, 6E) capacity */,
                         int64_t* element_edges /* (6,E) */, int8_t* signs /* (6,E) */, int64_t* nedges) {
    static const int le[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    int64_t cap = 1;
    while (cap < 12 * E + 16) cap <<= 1;
    int64_t* keys_a = malloc(sizeof(int64_t) * (size_t)cap);
    int64_t* keys_b = malloc(sizeof(int64_t) * (size_t)cap);
    int64_t* ids = malloc(sizeof(int64_t) * (size_t)cap);
    if (!keys_a || !keys_b || !ids) { free(keys_a); free(keys_b); free(ids); return fail("out of memory"); }
    for (int64_t k = 0; k < cap; ++k) ids[k] = -1;
    int64_t count = 0;
    for (int64_t e = 0; e < E; ++e) {
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j)
                if (t[i * E + e] == t[j * E + e]) {
                    free(keys_a); free(keys_b); free(ids);
                    snprintf(g_error, sizeof g_error, "cell %lld repeats a corner vertex", (long long)e);
                    return -2;
                }
        for (int l = 0; l < 6; ++l) {
            const int64_t f = t[le[l][0] * E + e], s = t[le[l][1] * E + e];
            const int64_t lo = f < s ? f : s, hi = f < s ? s : f;
            uint64_t h = (uint64_t)lo * 0x9E3779B97F4A7C15ull ^ ((uint64_t)hi + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
            int64_t slot = (int64_t)(h & (uint64_t)(cap - 1));
            while (ids[slot] >= 0 && !(keys_a[slot] == lo && keys_b[slot] == hi)) slot = (slot + 1) & (cap - 1);
            if (ids[slot] < 0) {
                ids[slot] = count; keys_a[slot] = lo; keys_b[slot] = hi;
                edges_out[count] = lo; edges_out[6 * E + count] = hi; /* rows stored at stride 6E */
                ++count;
            }
            element_edges[l * E + e] = ids[slot];
            signs[l * E + e] = f == lo ? 1 : -1;
        }
    }
    *nedges = count;
    free(keys_a); free(keys_b); free(ids);
    return 0;
}

/* integrate_tet_n1, edge_topology.cpp:156-233.  values/curls (E,6,3,Q). */
int oracle_integrate_tet_n1(const double* values, const double* curls, const double* dx,
                            const double* mass, const double* curl, int include_mass, int include_curl,
                            int64_t E, int64_t Q, double* out /* (E,6,6) */) {
    for (int64_t e = 0; e < E; ++e)
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                double res = 0.;
                for (int64_t q = 0; q < Q; ++q) {
                    double vd = 0., cd = 0.;
                    for (int i = 0; i < 3; ++i) {
                        const int64_t ro = ((e * 6 + r) * 3 + i) * Q + q, co = ((e * 6 + c) * 3 + i) * Q + q;
                        if (include_mass) vd += values[ro] * values[co];
                        if (include_curl) cd += curls[ro] * curls[co];
                    }
                    res += dx[e * Q + q] * (mass[e * Q + q] * vd + curl[e * Q + q] * cd);
                }
                out[(e * 6 + r) * 6 + c] = res;
            }
    return 0;
}

/* np.add.at(data, scatter.ravel(), elements.ravel()) after data.fill(0),
 * hcurl_assembler.py:94-97: serial, element-major order. */
int oracle_scatter_add(const int64_t* scatter, const double* elements, int64_t count, double* data, int64_t nnz) {
    memset(data, 0, sizeof(double) * (size_t)nnz);
    for (int64_t k = 0; k < count; ++k) data[scatter[k]] += elements[k];
    return 0;
}
