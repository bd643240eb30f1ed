"""Pins the solver / learned-FSAI oracle (oracle/fsai.py) by mathematics:
closed forms, optimality conditions, finite differences, brute force.
CPU only."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import fsai as of
from oracle import host, pipeline
import oracle


def p1_aniso_system(nx, ny, nz, seed, jitter=0.2):
    """Condensed P1 anisotropic system assembled by the C restatement."""
    rng = np.random.default_rng(seed)
    p, t = host.kuhn_mesh(np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1), np.linspace(0, 1, nz + 1))
    interior = np.all((p > 1e-12) & (p < 1 - 1e-12), axis=0)
    h = np.array([1 / nx, 1 / ny, 1 / nz])[:, None]
    p = p.copy()
    p[:, interior] += rng.uniform(-jitter, jitter, size=(3, interior.sum())) * h
    E = t.shape[1]
    g = rng.standard_normal((E, 3, 3))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    lam = 10. ** rng.uniform(-2, 2, size=(E, 3))
    A_e = np.einsum("eij,ej,ekj->eik", q, lam, q)
    X, W = host.tet_rule_4()
    (shape, grads, dx, x, det, tol, cond), _ = pipeline.basis_geometry(p, t, X, W, False)
    N, dofs = pipeline.dofs_of(t, p.shape[1], 4, 1)
    ip, ix, sc = oracle.pattern(dofs.reshape(E, -1), nrows=N)
    A_eq = np.ascontiguousarray(np.broadcast_to(A_e[:, None], (E, 4, 3, 3)))
    vals = oracle.assemble_tensor(grads, dx, A_eq, sc, len(ix))
    K = sp.csr_matrix((vals, ix, ip), shape=(N, N))
    b = np.zeros(N)
    np.add.at(b, t.T.ravel(), np.repeat(dx.sum(1) / 4, 4))
    Af, bf, free = of.condense(K, b, interior)
    return Af, bf, p[:, free].T.copy()


@pytest.fixture(scope="module")
def small():
    return p1_aniso_system(4, 5, 3, 7)


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert of.bf16(np.float32(1. + 2. ** -8)) == 1.0
    assert of.bf16(np.float32(1. + 3 * 2. ** -8)) == 1. + 2 ** -6
    assert of.bf16(np.float32(-2.5)) == -2.5
    x = np.float32(np.pi)
    assert abs(of.bf16(x) - np.pi) <= 2. ** -8 * np.pi


def test_condense_matches_setdiff_definition(small):
    A, b, _ = small
    keep = np.ones(A.shape[0], dtype=bool)
    keep[[0, 3]] = False
    Ac, bc, free = of.condense(A, b, keep)
    free_ref = np.setdiff1d(np.arange(A.shape[0]), [0, 3])
    np.testing.assert_array_equal(free, free_ref)
    np.testing.assert_array_equal(Ac.toarray(), A.toarray()[np.ix_(free_ref, free_ref)])


def test_kaporin_identity_closed_form(small):
    A, _, _ = small
    lptr, lidx = of.lower_pattern(A)
    o = np.zeros(len(lidx))
    loss, grad = of.kaporin(A, [0, A.shape[0]], lptr, lidx, o)
    # Gt = I: tr(At) = n and log det = 0
    assert abs(loss[0]) < 1e-14
    d = A.diagonal()
    rows = np.repeat(np.arange(A.shape[0]), np.diff(lptr))
    at = np.asarray(A[rows, lidx]).ravel() / np.sqrt(d[rows] * d[lidx])
    n = A.shape[0]
    expect = np.where(rows == lidx, 0., 2. * at / n)
    np.testing.assert_allclose(grad, expect, atol=1e-15)


def test_kaporin_zero_gradient_at_exact_fsai(small):
    A, _, _ = small
    n = A.shape[0]
    rs = 1. / np.sqrt(A.diagonal())
    At = sp.csr_matrix(sp.diags(rs) @ A @ sp.diags(rs))
    G = of.exact_fsai(At)
    lptr, lidx = of.lower_pattern(A)
    G.sort_indices()
    np.testing.assert_array_equal(G.indptr, lptr)
    g = G.data.copy()
    rows = np.repeat(np.arange(n), np.diff(lptr))
    diag = rows == lidx
    o = g.copy()
    o[diag] = np.log(g[diag])
    loss, grad = of.kaporin(A, [0, n], lptr, lidx, o)
    assert np.max(np.abs(grad)) < 1e-10
    # the optimum has unit diagonal of G At G^T
    np.testing.assert_allclose((G @ At @ G.T).diagonal(), 1., rtol=1e-12)
    # and any perturbation increases the loss
    rng = np.random.default_rng(0)
    for _ in range(3):
        l2, _ = of.kaporin(A, [0, n], lptr, lidx, o + 1e-3 * rng.standard_normal(len(o)))
        assert l2[0] > loss[0]


def test_kaporin_scale_invariant_and_finite_differences(small):
    A, _, _ = small
    n = A.shape[0]
    half = n // 2
    # two "systems": split the matrix into two diagonal blocks
    Ab = sp.block_diag([A[:half, :half], A[half:, half:]]).tocsr()
    sys_ptr = [0, half, n]
    lptr, lidx = of.lower_pattern(Ab)
    rng = np.random.default_rng(1)
    rows = np.repeat(np.arange(n), np.diff(lptr))
    o = np.where(rows == lidx, 0.1 * rng.standard_normal(len(lidx)), -0.2 + 0.1 * rng.standard_normal(len(lidx)))
    loss, grad = of.kaporin(Ab, sys_ptr, lptr, lidx, o)
    # scaling Gt by c: off-diagonal o -> c o, diagonal o -> o + log c
    c = 3.7
    o2 = np.where(rows == lidx, o + np.log(c), c * o)
    loss2, _ = of.kaporin(Ab, sys_ptr, lptr, lidx, o2)
    np.testing.assert_allclose(loss2, loss, rtol=1e-12, atol=1e-13)
    eps = 1e-6
    for q in rng.choice(len(o), size=12, replace=False):
        e = np.zeros_like(o)
        e[q] = eps
        lp, _ = of.kaporin(Ab, sys_ptr, lptr, lidx, o + e)
        lm, _ = of.kaporin(Ab, sys_ptr, lptr, lidx, o - e)
        fd = (lp.sum() - lm.sum()) / (2 * eps)
        assert abs(fd - grad[q]) < 1e-7 * max(1., abs(grad[q]))


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    flat = rng.standard_normal(of.NPARAMS) * 0.2
    X = rng.standard_normal((40, of.NF))
    go = rng.standard_normal(40)
    grad = of.mlp_backward(flat, X, go, rounding=False)
    eps = 1e-6
    blocks = {"W1": 0, "b1": 2048, "W2": 2176, "b2": 18560, "w3": 18688, "b3": 18816}
    picks = [blocks[k] + j for k in blocks for j in ((0,) if k == "b3" else (0, 5))]
    picks += [blocks["W2"] + 1000, blocks["W1"] + 777]
    for q in picks:
        e = np.zeros_like(flat)
        e[q] = eps
        fp = of.mlp_forward(flat + e, X, rounding=False)[0] @ go
        fm = of.mlp_forward(flat - e, X, rounding=False)[0] @ go
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - grad[q]) < 1e-6 * max(1., abs(grad[q])), q


def test_mlp_zero_output_layer_is_zero():
    flat = np.zeros(of.NPARAMS)
    flat[:2048] = 1.
    o, _ = of.mlp_forward(flat, np.ones((3, of.NF)))
    np.testing.assert_array_equal(o, 0.)


def test_features_brute_force(small):
    A, _, xyz = small
    lptr, lidx, F = of.features(A, xyz)
    D = A.toarray()
    d = np.diag(D)
    At = D / np.sqrt(np.outer(d, d))
    n = len(d)
    rng = np.random.default_rng(5)
    for q in rng.choice(len(lidx), size=25, replace=False):
        i = int(np.searchsorted(lptr, q, side="right") - 1)
        j = int(lidx[q])
        assert F[q, 0] == pytest.approx(At[i, j], rel=1e-14)
        assert F[q, 1] == (1. if i == j else 0.)
        assert F[q, 2] == pytest.approx(np.abs(At[i]).sum() - 1., rel=1e-12)
        assert F[q, 4] == pytest.approx(D[i].sum() / d[i], rel=1e-10, abs=1e-12)
        ks = np.arange(n)
        assert F[q, 10] == pytest.approx(At[i] @ At[:, j], rel=1e-12)
        assert F[q, 9] == pytest.approx(At[i, ks < j] @ At[j, ks < j], rel=1e-12, abs=1e-15)
        m = (ks > j) & (ks < i)
        assert F[q, 8] == pytest.approx(At[i, m] @ At[m, j], rel=1e-12, abs=1e-15)
        assert F[q, 11] == pytest.approx(0.5 * np.log(d[i] / d[j]), abs=1e-14)


def test_pcg_oracle_solution_and_diagonal_case(small):
    A, b, _ = small
    d = A.diagonal()
    x, it = of.pcg(A, b, lambda r: r / d, tol=1e-12)
    np.testing.assert_allclose(x, spla.spsolve(A.tocsc(), b), rtol=1e-9)
    # CG on a diagonal matrix with 3 distinct eigenvalues converges in 3 steps
    Dm = sp.diags(np.repeat([1., 2., 5.], 10)).tocsr()
    _, it3 = of.pcg(Dm, np.ones(30), lambda r: r, tol=1e-12)
    assert it3 == 3
    # Jacobi on a diagonal matrix: one step
    _, it1 = of.pcg(Dm, np.ones(30), lambda r: r / Dm.diagonal(), tol=1e-12)
    assert it1 == 1


def test_exact_fsai_beats_jacobi(small):
    A, b, _ = small
    d = A.diagonal()
    _, itj = of.pcg(A, b, lambda r: r / d)
    G = of.exact_fsai(A)
    _, itf = of.pcg(A, b, lambda r: G.T @ (G @ r))
    assert itf < itj
