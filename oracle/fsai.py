"""CPU oracle for the batched solver and the learned FSAI -- TEST
INFRASTRUCTURE ONLY (imported by tests/ and nothing shipped).

These parts have no counterpart in the reference (its solver policy is
external, python/skfemntv/capabilities.py:67); the only reference anchor is
the condensed solve of examples/scikit-fem-gallery/comparison.py:80-87, which
``condense`` restates.  Every function here is pinned by mathematics in
tests/test_fsai_oracle.py rather than by reference vectors:

* ``kaporin``         closed form at G = I / exact FSAI optimum (zero
                      gradient), scale invariance, finite differences;
* ``mlp_forward`` /
  ``mlp_backward``    finite differences of the unrounded network, and
                      the BF16 rounding points shared with the kernels;
* ``features``        brute-force dense recomputation on tiny systems;
* ``pcg``             scipy's direct solve (same solution) and the known
                      iteration count of CG on a diagonal matrix;
* ``exact_fsai``      Kaporin optimality (zero gradient of ``kaporin``).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

NF = 16
H = 128
SIZES = {"W1": (NF, H), "b1": (H,), "W2": (H, H), "b2": (H,), "w3": (H,), "b3": (1,)}
ORDER = ("W1", "b1", "W2", "b2", "w3", "b3")
NPARAMS = sum(int(np.prod(s)) for s in SIZES.values())


def unflatten(flat):
    out, o = {}, 0
    for k in ORDER:
        n = int(np.prod(SIZES[k]))
        out[k] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(SIZES[k])
        o += n
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[k], dtype=np.float64).ravel() for k in ORDER])


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def condense(A, b, keep):
    """A[free][:, free], b[free] with free = flatnonzero(keep)
    (examples/scikit-fem-gallery/comparison.py:80-87)."""
    free = np.flatnonzero(keep)
    A = sp.csr_matrix(A)
    return A[free][:, free].tocsr(), np.asarray(b)[free], free


def lower_pattern(A):
    L = sp.tril(sp.csr_matrix(A), format="csr")
    L.sort_indices()
    return L.indptr.astype(np.int64), L.indices.astype(np.int64)


def features(A, xyz=None):
    """(lptr, lidx, F (nnz_lower, 16) float64) -- see the kernel's list."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    d = A.diagonal()
    rs = 1. / np.sqrt(d)
    At = sp.diags(rs) @ A @ sp.diags(rs)
    At = sp.csr_matrix(At)
    lptr, lidx = lower_pattern(A)
    off = At - sp.diags(At.diagonal())
    abs_sum = np.asarray(abs(off).sum(axis=1)).ravel()
    rowsum = np.asarray(A.sum(axis=1)).ravel() / d
    rowlen = np.diff(A.indptr) / 16.
    h = np.ones(n)
    if xyz is not None:
        for i in range(n):
            cols = A.indices[A.indptr[i]:A.indptr[i + 1]]
            cols = cols[cols != i]
            if len(cols):
                h[i] = np.mean(np.linalg.norm(xyz[cols] - xyz[i], axis=1))
    rows = [dict(zip(At.indices[At.indptr[i]:At.indptr[i + 1]], At.data[At.indptr[i]:At.indptr[i + 1]]))
            for i in range(n)]
    F = np.zeros((len(lidx), NF))
    for i in range(n):
        for q in range(lptr[i], lptr[i + 1]):
            j = int(lidx[q])
            ri, rj = rows[i], rows[j]
            common = set(ri) & set(rj)
            t_all = sum(ri[k] * rj[k] for k in common)
            t_lo = sum(ri[k] * rj[k] for k in common if k < j)
            t_mid = sum(ri[k] * rj[k] for k in common if j < k < i)
            f = F[q]
            f[0] = ri.get(j, 0.)
            f[1] = 1. if i == j else 0.
            f[2], f[3] = abs_sum[i], abs_sum[j]
            f[4], f[5] = rowsum[i], rowsum[j]
            f[6], f[7] = rowlen[i], rowlen[j]
            f[8], f[9], f[10] = t_mid, t_lo, t_all
            f[11] = 0.5 * np.log(d[i] / d[j])
            if xyz is not None and i != j:
                v = (xyz[j] - xyz[i]) / h[i]
                f[12:15] = v
                f[15] = np.linalg.norm(v)
    return lptr, lidx, F


def gtilde_from_outputs(o, lptr, lidx):
    """Gt values: o off the diagonal, exp(o) on it (o clamped to [-30, 30])."""
    g = np.asarray(o, dtype=np.float64).copy()
    n = len(lptr) - 1
    rows = np.repeat(np.arange(n), np.diff(lptr))
    diag = rows == lidx
    g[diag] = np.exp(np.clip(g[diag], -30., 30.))
    return g, rows, diag


def kaporin(A, sys_ptr, lptr, lidx, o):
    """Per-system loss log(tr(Gt At Gt^T)/n) - (2/n) sum log Gt_ii and its
    gradient with respect to the network outputs o (per lower nonzero)."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    rs = 1. / np.sqrt(A.diagonal())
    At = sp.csr_matrix(sp.diags(rs) @ A @ sp.diags(rs))
    g, rows, diag = gtilde_from_outputs(o, lptr, lidx)
    G = sp.csr_matrix((g, lidx, lptr), shape=(n, n))
    GA = (G @ At).tocsr()
    P = np.asarray(GA[rows, lidx]).ravel()
    S = len(sys_ptr) - 1
    loss = np.zeros(S)
    grad = np.zeros(len(lidx))
    for s in range(S):
        r0, r1 = sys_ptr[s], sys_ptr[s + 1]
        q0, q1 = lptr[r0], lptr[r1]
        ns = r1 - r0
        tr = np.sum(g[q0:q1] * P[q0:q1])
        logd = np.sum(np.log(g[q0:q1][diag[q0:q1]]))
        loss[s] = np.log(tr / ns) - 2. * logd / ns
        gr = 2. * P[q0:q1] / tr
        dg = diag[q0:q1]
        gr[dg] = gr[dg] * g[q0:q1][dg] - 2. / ns
        grad[q0:q1] = gr
    return loss, grad


def factor_values(o, lptr, lidx, A):
    """G = Gt D^{-1/2} values."""
    g, _, _ = gtilde_from_outputs(o, lptr, lidx)
    rs = 1. / np.sqrt(sp.csr_matrix(A).diagonal())
    return g * rs[lidx]


def mlp_forward(flat, X, rounding=True):
    """Network outputs with the kernels' BF16 rounding points (or none)."""
    p = unflatten(flat)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    x = r(X)
    z1 = x @ r(p["W1"]) + p["b1"]
    h1 = r(np.maximum(z1, 0.))
    z2 = h1 @ r(p["W2"]) + p["b2"]
    h2 = np.maximum(z2, 0.)
    o = h2 @ p["w3"] + p["b3"][0]
    return o, (x, z1, h1, z2, h2)


def mlp_backward(flat, X, go, rounding=True):
    """Gradient of sum_k go[k] * o[k] with respect to the flat parameters,
    with the kernels' BF16 rounding of dz2, dz1 and the weight operands."""
    p = unflatten(flat)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    _, (x, z1, h1, z2, h2) = mlp_forward(flat, X, rounding)
    go = np.asarray(go, dtype=np.float64)
    g = {}
    g["b3"] = np.array([go.sum()])
    g["w3"] = h2.T @ go
    dz2 = np.where(z2 > 0., go[:, None] * p["w3"][None, :], 0.)
    g["b2"] = dz2.sum(axis=0)
    dz2r = r(dz2)
    g["W2"] = h1.T @ dz2r
    dh1 = dz2r @ r(p["W2"]).T
    dz1 = np.where(z1 > 0., dh1, 0.)
    g["b1"] = dz1.sum(axis=0)
    g["W1"] = x.T @ r(dz1)
    return flatten(g)


def pcg(A, b, apply_minv, tol=1e-8, max_iter=20000):
    """Textbook PCG from x0 = 0 with the same stopping rule as the kernels
    (recursive residual ||r|| <= tol ||b||).  Returns (x, iterations)."""
    A = sp.csr_matrix(A)
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_minv(r)
    p = z.copy()
    rz = r @ z
    bb = b @ b
    for it in range(1, max_iter + 1):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = apply_minv(r)
        rz_new = r @ z
        if not (r @ r > tol * tol * bb):
            return x, it
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_iter


def exact_fsai(A):
    """Classical FSAI on A's lower pattern: row i solves A[S,S] g = e_i and is
    scaled so that (G A G^T)_ii = 1 (the Kaporin optimum for that pattern)."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    lptr, lidx = lower_pattern(A)
    vals = np.zeros(len(lidx))
    for i in range(n):
        S = lidx[lptr[i]:lptr[i + 1]]
        sub = A[S][:, S].toarray()
        e = (S == i).astype(float)
        g = np.linalg.solve(sub, e)
        g /= np.sqrt(g[S == i][0])
        vals[lptr[i]:lptr[i + 1]] = g
    return sp.csr_matrix((vals, lidx, lptr), shape=(n, n))
