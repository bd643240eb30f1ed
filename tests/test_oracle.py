"""The CPU oracle is pinned before it is trusted (no GPU needed).

Three kinds of checks:
* golden vectors written from the reference itself (tests/golden/fixtures.npz);
* the reference's own compiled C++ (oracle/_ref) on identical inputs, where built;
* closed forms / invariants chosen so that a dropped term, a sign or index
  error, or a transposed operand fails.
"""

import os

import numpy as np
import pytest

import oracle
from oracle import host, pipeline

FIX = os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(FIX, allow_pickle=False)


@pytest.fixture(scope="module")
def ref():
    r = oracle.ref_module()
    if r is None:
        pytest.skip("oracle/_ref (reference C++ build) not present")
    return r


def setup_module(module):
    if not os.path.exists(os.path.join(os.path.dirname(oracle.__file__), "liboracle.so")):
        oracle.build()


def rule4():
    return host.tet_rule_4()


# ------------------------------------------------------------------ golden
def test_mesh_numbering_matches_reference_golden(golden):
    p, t = host.kuhn_mesh(np.linspace(0, 1, 3), np.linspace(0, 1, 4), np.linspace(0, 1, 3))
    np.testing.assert_array_equal(p, golden["kuhn_p"])
    np.testing.assert_array_equal(t, golden["kuhn_t"])
    p2, t2 = host.p2_mesh(p, t)
    np.testing.assert_array_equal(p2, golden["p2_p"])
    np.testing.assert_array_equal(t2, golden["p2_t"])
    for tag, pp, tt, nodes, dim in (("p1", p, t, 4, 1), ("p2", p2, t2, 10, 1), ("vp2", p2, t2, 10, 3)):
        N, ed = host.element_dofs(tt[:nodes], pp.shape[1], dim)
        assert N == int(golden[f"dofs_{tag}_N"])
        np.testing.assert_array_equal(ed, golden[f"dofs_{tag}"])


def test_c1_p1_laplace_and_load_match_golden(golden):
    p, t = golden["c1_p"], golden["c1_t"]
    X, W = rule4()
    E = t.shape[1]
    ip, ix, vals, N = pipeline.bilinear(p, t, X, W, False, gradient=np.ones((E, 4)))
    np.testing.assert_array_equal(ip, golden["c1_indptr"])
    np.testing.assert_array_equal(ix, golden["c1_indices"])
    assert pipeline.relative_frobenius(vals, golden["c1_data"]) <= 1e-14
    rhs = pipeline.load_vector(p, t, X, W, False, [1.])
    assert pipeline.relative_frobenius(rhs, golden["c1_rhs"]) <= 1e-14


def _jit(golden):
    p, t = golden["jit_p"], golden["jit_t"]
    p2, t2 = host.p2_mesh(p, t)
    return p, t, p2, t2


def test_c2_p2_anisotropic_matches_golden(golden):
    p, t, p2, t2 = _jit(golden)
    X, W = rule4()
    E = t.shape[1]
    A = golden["c2_A"]
    A_eq = host.anisotropic_tensor(np.broadcast_to(np.moveaxis(A, 0, -1)[..., None], (3, 3, E, 4)), (E, 4))
    ip, ix, vals, N = pipeline.tensor(p2, t2, X, W, True, A_eq)
    np.testing.assert_array_equal(ip, golden["c2_indptr"])
    np.testing.assert_array_equal(ix, golden["c2_indices"])
    assert pipeline.relative_frobenius(vals, golden["c2_data"]) <= 1e-14


def test_c2_nonsymmetric_native_layout_matches_golden(golden):
    p, t, p2, t2 = _jit(golden)
    X, W = rule4()
    x = np.moveaxis(golden["c2n_x"], -1, 0)
    E = t.shape[1]
    T = np.empty((E, 4, 3, 3))
    T[..., 0, 0] = 1. + x[0]; T[..., 0, 1] = 0.3 + 0.1 * x[1]; T[..., 0, 2] = 0.
    T[..., 1, 0] = -0.2; T[..., 1, 1] = 2. - 0.4 * x[0]; T[..., 1, 2] = 0.1 * x[2]
    T[..., 2, 0] = 0.05; T[..., 2, 1] = 0.; T[..., 2, 2] = 1.5
    ip, ix, vals, N = pipeline.tensor(p2, t2, X, W, True, host.anisotropic_tensor(T, (E, 4)))
    np.testing.assert_array_equal(ix, golden["c2n_indices"])
    assert pipeline.relative_frobenius(vals, golden["c2n_data"]) <= 1e-14


def test_c3_vector_p2_elasticity_matches_golden(golden):
    p, t, p2, t2 = _jit(golden)
    X, W = rule4()
    mu, lam = golden["c3_mu"], golden["c3_lam"]
    ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, components=3, symgrad=np.repeat(2 * mu[:, None], 4, 1),
                                        divergence=np.repeat(lam[:, None], 4, 1))
    np.testing.assert_array_equal(ip, golden["c3_indptr"])
    np.testing.assert_array_equal(ix, golden["c3_indices"])
    assert pipeline.relative_frobenius(vals, golden["c3_data"]) <= 1e-14


def test_c4_n1_maxwell_matches_golden(golden):
    p, t = golden["jit_p"], golden["jit_t"]
    E = t.shape[1]
    X, W = host.tet_rule_duffy(2)
    np.testing.assert_array_equal(X, golden["c4_X"])
    np.testing.assert_array_equal(W, golden["c4_W"])
    kappa = golden["c4_kappa"]
    Q = len(W)
    ip, ix, data, edges, ee, signs, _ = pipeline.n1_maxwell(p, t, np.repeat(kappa[:, None], Q, 1), np.ones((E, Q)))
    np.testing.assert_array_equal(edges, golden["c4_edges"])
    np.testing.assert_array_equal(ee, golden["c4_element_edges"])
    np.testing.assert_array_equal(signs, golden["c4_signs"])
    np.testing.assert_array_equal(ip, golden["c4_indptr"])
    np.testing.assert_array_equal(ix, golden["c4_indices"])
    assert pipeline.relative_frobenius(data, golden["c4_data"]) <= 1e-14
    dx, values, curls, det = host.n1_basis(p, t, X, W, signs)
    np.testing.assert_allclose(values[0], golden["c4_values0"], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(curls[0], golden["c4_curls0"], rtol=1e-14, atol=1e-14)


# ------------------------------------------------------------------ reference C++ (oracle/_ref)
def _small_p2(seed=1, n=3):
    ax = np.linspace(0, 1, n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, 1. / n, 0.2, seed)
    return host.p2_mesh(p, t)


def test_c_geometry_and_pattern_are_bit_exact_vs_reference_cpp(ref):
    p2, t2 = _small_p2()
    X, W = host.tet_rule_keast11()
    sh, rg = host.tabulate_p2(X)
    mine, bad = oracle.tabulate_geometry(p2, t2, sh, rg, sh, rg, W)
    theirs = ref.tabulate_basis_geometry(p2, t2, sh, rg, sh, rg, W)
    for a, b in zip(mine, theirs):
        np.testing.assert_array_equal(a, np.asarray(b))
    assert bad == (-1, -1, 0)
    N, dofs = pipeline.dofs_of(t2, p2.shape[1], 10, 3)
    E = t2.shape[1]
    ip, ix, sc = oracle.pattern(dofs.reshape(E, -1), nrows=N)
    B = ref.BilinearFormAssembler(dofs, mine[0], mine[1], mine[2])
    np.testing.assert_array_equal(ip, B.indptr)
    np.testing.assert_array_equal(ix, B.indices)


@pytest.mark.parametrize("kinds", [("value",), ("gradient",), ("symgrad",), ("divergence",),
                                   ("value", "gradient", "symgrad", "divergence")])
@pytest.mark.parametrize("components", [1, 3])
def test_c_bilinear_kernel_matches_reference_cpp(ref, kinds, components):
    p2, t2 = _small_p2(seed=components)
    X, W = rule4()
    sh, rg = host.tabulate_p2(X)
    (shape, grads, dx, *_), _ = oracle.tabulate_geometry(p2, t2, sh, rg, sh, rg, W)
    N, dofs = pipeline.dofs_of(t2, p2.shape[1], 10, components)
    E = t2.shape[1]
    rng = np.random.default_rng(len(kinds))
    coef = {k: rng.uniform(0.5, 2.0, (E, 4)) for k in kinds}
    B = ref.BilinearFormAssembler(dofs, shape, grads, dx)
    theirs, _ = B.assemble(coef.get("value"), coef.get("gradient"), coef.get("symgrad"), coef.get("divergence"), 1)
    ip, ix, sc = oracle.pattern(dofs.reshape(E, -1), nrows=N)
    mine = oracle.assemble_bilinear(shape, grads, dx, sc, len(ix), components, value=coef.get("value"),
                                    gradient=coef.get("gradient"), symgrad=coef.get("symgrad"),
                                    divergence=coef.get("divergence"))
    np.testing.assert_array_equal(mine, theirs)
    colour, ncol = oracle.colouring(dofs.reshape(E, -1), N)
    threaded = oracle.assemble_bilinear(shape, grads, dx, sc, len(ix), components, value=coef.get("value"),
                                        gradient=coef.get("gradient"), symgrad=coef.get("symgrad"),
                                        divergence=coef.get("divergence"), threads=4, colour=colour)
    assert pipeline.relative_frobenius(threaded, theirs) <= 1e-15


def test_c_tensor_and_linear_kernels_match_reference_cpp(ref):
    p2, t2 = _small_p2(seed=4)
    X, W = rule4()
    sh, rg = host.tabulate_p2(X)
    (shape, grads, dx, *_), _ = oracle.tabulate_geometry(p2, t2, sh, rg, sh, rg, W)
    N, dofs = pipeline.dofs_of(t2, p2.shape[1], 10, 1)
    E = t2.shape[1]
    A = np.random.default_rng(0).standard_normal((E, 4, 3, 3))
    Cb = ref.CrossBilinearAssembler(dofs, dofs, shape, shape, dx, grads, grads)
    theirs, _ = Cb.assemble(np.ascontiguousarray(A[:, :, None, :, None, :]), "gradient", "gradient", 1)
    ip, ix, sc = oracle.pattern(dofs.reshape(E, -1), nrows=N)
    np.testing.assert_array_equal(oracle.assemble_tensor(grads, dx, A, sc, len(ix)), theirs)
    Lf = ref.LinearFormAssembler(dofs, shape, grads, dx)
    f = np.random.default_rng(1).standard_normal((E, 4, 1))
    G = np.random.default_rng(2).standard_normal((E, 4, 1, 3))
    theirs, _ = Lf.assemble(f, G, 1)
    np.testing.assert_array_equal(oracle.assemble_linear(shape, grads, dx, dofs, N, 1, value=f, gradient=G), theirs)
    theirs, _ = Lf.assemble(np.array([1.]), None, 1)
    np.testing.assert_array_equal(oracle.assemble_linear(shape, grads, dx, dofs, N, 1, value=np.array([1.])), theirs)


def test_c_edge_topology_n1_integration_scatter_match_reference_cpp(ref):
    ax = np.linspace(0, 1, 4)
    p, t = host.kuhn_mesh(ax, ax, ax)
    t = t.copy()
    sel = np.arange(t.shape[1]) % 3 == 0  # mixed orientation
    t[[0, 1]] = np.where(sel, t[[1, 0]], t[[0, 1]])
    edges, ee, signs = oracle.edge_topology(t)
    re, ree, rs = ref.build_oriented_edge_topology(t, 3)
    np.testing.assert_array_equal(edges, re)
    np.testing.assert_array_equal(ee, ree)
    np.testing.assert_array_equal(signs, rs)
    ip, ix, sc = oracle.pattern(ee.T.copy(), nrows=edges.shape[1])
    rip, rix, rsc = ref.build_edge_csr_pattern(ee, edges.shape[1])
    np.testing.assert_array_equal(ip, rip)
    np.testing.assert_array_equal(ix, rix)
    np.testing.assert_array_equal(sc, rsc)
    X, W = host.tet_rule_duffy(2)
    dx, values, curls, det = host.n1_basis(p, t, X, W, signs)
    E, Q = t.shape[1], len(W)
    k = np.random.default_rng(3).uniform(0.1, 10, (E, Q))
    c = np.random.default_rng(4).uniform(0.1, 10, (E, Q))
    mine = oracle.integrate_tet_n1(values, curls, dx, k, c)
    theirs = np.empty((E, 6, 6))
    ref.integrate_tet_n1(values, curls, dx, k, c, True, True, theirs)
    np.testing.assert_array_equal(mine, theirs)
    data = np.zeros(len(ix))
    np.add.at(data, sc.ravel(), mine.ravel())
    np.testing.assert_array_equal(oracle.scatter_add(sc, mine, len(ix)), data)


def test_edge_topology_rejects_repeated_corner():
    t = np.array([[0], [0], [1], [2]], dtype=np.int64)
    with pytest.raises(ValueError, match="cell 0 repeats a corner vertex"):
        oracle.edge_topology(t)


# ------------------------------------------------------------------ closed forms
@pytest.mark.parametrize("rule,degree", [(host.tet_rule_4, 2), (host.tet_rule_keast11, 4),
                                         (lambda: host.tet_rule_duffy(2), 3)])
def test_quadrature_integrates_monomials_exactly(rule, degree):
    from math import factorial
    X, W = rule()
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            for c in range(degree + 1 - a - b):
                exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                got = np.sum(W * X[0] ** a * X[1] ** b * X[2] ** c)
                assert abs(got - exact) <= 1e-15, (a, b, c)


def test_tabulation_closed_forms():
    X, W = host.tet_rule_keast11()
    for tab in (host.tabulate_p1, host.tabulate_p2):
        s, g = tab(X)
        np.testing.assert_allclose(s.sum(axis=1), 1., atol=1e-14)
        np.testing.assert_allclose(g.sum(axis=1), 0., atol=1e-14)
    # Kronecker property at the P2 nodes and finite-difference gradients
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [.5, .5, 0], [0, .5, 0],
                      [0, 0, .5], [.5, 0, .5], [0, .5, .5]]).T
    s, _ = host.tabulate_p2(nodes)
    np.testing.assert_allclose(s, np.eye(10), atol=1e-15)
    pt = np.array([[.2], [.3], [.1]])
    _, g = host.tabulate_p2(pt)
    h = 1e-6
    for j in range(3):
        e = np.zeros((3, 1)); e[j] = h
        fd = (host.tabulate_p2(pt + e)[0] - host.tabulate_p2(pt - e)[0]) / (2 * h)
        np.testing.assert_allclose(g[0, :, j], fd[0], atol=1e-8)


def test_p1_kuhn_laplacian_closed_form():
    n = 4
    ax = np.linspace(0, 1, n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    X, W = rule4()
    E = t.shape[1]
    ip, ix, vals, N = pipeline.bilinear(p, t, X, W, False, gradient=np.ones((E, 4)))
    from scipy.sparse import csr_matrix
    K = csr_matrix((vals, ix, ip), shape=(N, N))
    h = 1. / n
    c = 2 + n + 1 + 2 * (n + 1) ** 2  # interior vertex (2,1,2)
    row = K.getrow(c).toarray().ravel()
    nz = np.flatnonzero(np.abs(row) > 1e-12)
    assert abs(row[c] - 6 * h) < 1e-14
    np.testing.assert_allclose(row[nz[nz != c]], -h, atol=1e-14)
    assert len(nz) == 7  # the pattern keeps exact-zero face/body diagonals, the values vanish
    assert ip[c + 1] - ip[c] == 15
    np.testing.assert_allclose(K @ np.ones(N), 0., atol=1e-13)
    for i in range(3):
        for j in range(3):
            assert abs(p[i] @ (K @ p[j]) - (1. if i == j else 0.)) < 1e-13


@pytest.mark.parametrize("quadratic", [False, True])
def test_anisotropic_energy_of_linear_fields(quadratic):
    """e_l^T K e_k = sum_e |T_e| A_e[l,k] for the interpolants of x_l, x_k;
    nonsymmetric A pins row = test orientation."""
    ax = np.linspace(0, 1, 3)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, .5, 0.2, 3)
    if quadratic:
        p, t = host.p2_mesh(p, t)
    X, W = rule4()
    E = t.shape[1]
    A = np.random.default_rng(0).standard_normal((E, 3, 3))
    A_eq = np.ascontiguousarray(np.broadcast_to(A[:, None], (E, 4, 3, 3)))
    ip, ix, vals, N = pipeline.tensor(p, t, X, W, quadratic, A_eq)
    from scipy.sparse import csr_matrix
    K = csr_matrix((vals, ix, ip), shape=(N, N))
    v = p[:, t[:4]]
    vol = np.abs(np.linalg.det(np.stack([v[:, k] - v[:, 0] for k in (1, 2, 3)], axis=1).transpose(2, 0, 1))) / 6
    for l in range(3):
        for k in range(3):
            expect = np.sum(vol * A[:, l, k])
            assert abs(p[l] @ (K @ p[k]) - expect) <= 1e-12 * max(1., abs(expect))


def test_elasticity_rigid_modes_and_symmetry():
    ax = np.linspace(0, 1, 3)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, .5, 0.15, 5)
    p2, t2 = host.p2_mesh(p, t)
    X, W = rule4()
    E = t.shape[1]
    rng = np.random.default_rng(1)
    mu = rng.uniform(.5, 2, E)
    lam = rng.uniform(.5, 2, E)
    ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, components=3, symgrad=np.repeat(2 * mu[:, None], 4, 1),
                                        divergence=np.repeat(lam[:, None], 4, 1))
    from scipy.sparse import csr_matrix
    K = csr_matrix((vals, ix, ip), shape=(N, N))
    assert abs(K - K.T).max() <= 1e-13 * abs(K).max()
    x = p2[:, np.arange(N // 3)]
    modes = []
    for c in range(3):
        r = np.zeros((N // 3, 3)); r[:, c] = 1.; modes.append(r.ravel())
    for (i, j) in ((0, 1), (1, 2), (0, 2)):
        r = np.zeros((N // 3, 3)); r[:, i] = -x[j]; r[:, j] = x[i]; modes.append(r.ravel())
    for r in modes:
        assert np.abs(K @ r).max() <= 1e-12 * abs(K).max()


def test_n1_reference_moments_and_curl_kernel():
    # edge moments of the reference Whitney functions are the identity
    g, gw = np.polynomial.legendre.leggauss(4)
    s = .5 * (g + 1)
    gw = .5 * gw
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.]])
    M = np.zeros((6, 6))
    for k, (a, b) in enumerate(host.TET_EDGES):
        tang = V[b] - V[a]
        pts = V[a][:, None] + tang[:, None] * s[None, :]
        vals, _ = host.n1_reference(pts)
        M[k] = np.einsum("biq,i,q->b", vals, tang, gw)
    np.testing.assert_allclose(M, np.eye(6), atol=2e-15)
    # curl-curl annihilates discrete gradients; constant fields give |c|^2 kappa vol
    ax = np.linspace(0, 1, 3)
    p, t = host.kuhn_mesh(ax, ax, ax)
    p = host.jitter_interior(p, .5, 0.1, 1)
    E = t.shape[1]
    ip, ix, data, edges, ee, signs, (X, W) = pipeline.n1_maxwell(p, t, np.zeros((E, 27)), np.ones((E, 27)),
                                                               include_mass=False)
    from scipy.sparse import csr_matrix
    Kc = csr_matrix((data, ix, ip), shape=(edges.shape[1],) * 2)
    phi = np.random.default_rng(0).standard_normal(p.shape[1])
    grad_dofs = phi[edges[1]] - phi[edges[0]]
    assert np.abs(Kc @ grad_dofs).max() <= 1e-12 * abs(Kc).max()
    kappa = np.random.default_rng(2).uniform(.5, 2, E)
    ip, ix, data, *_ = pipeline.n1_maxwell(p, t, np.repeat(kappa[:, None], 27, 1), np.zeros((E, 27)), include_curl=False)
    Km = csr_matrix((data, ix, ip), shape=(edges.shape[1],) * 2)
    cvec = np.array([1.25, -.5, .75])
    u = (p[:, edges[1]] - p[:, edges[0]]).T @ cvec
    v = p[:, t]
    vol = np.abs(np.linalg.det(np.stack([v[:, k] - v[:, 0] for k in (1, 2, 3)], axis=1).transpose(2, 0, 1))) / 6
    assert abs(u @ (Km @ u) - np.sum(kappa * vol) * cvec @ cvec) <= 1e-12


# ------------------------------------------------------------------ cross blocks
# (row field, column field, row kind, column kind, coefficient form); fields
# are (quadratic, components); row = test, column = trial
_CROSS_CASES = {
    "column_divergence_p1_vp2": ((False, 1), (True, 3), "value", "gradient", "coldiv"),
    "row_divergence_vp2_p1": ((True, 3), (False, 1), "gradient", "value", "rowdiv"),
    "mass_p1_p2": ((False, 1), (True, 1), "value", "value", "scalar"),
    "gradient_vp2_vp1": ((True, 3), (False, 3), "gradient", "gradient", "scalar"),
    "tensor_vp1_p2": ((False, 3), (True, 1), "gradient", "value", "tensor"),
}


def _cross_inputs(case, p2, t2, X, W, seed):
    (rq, Cr), (cq, Cc), rk, ck, form = _CROSS_CASES[case]
    E, Q = t2.shape[1], len(W)
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.5, 2.0, (E, Q))
    if form == "scalar":
        return c, False
    if form == "coldiv":
        T = np.zeros((E, Q, 1, 3, 3))
        for k in range(3):
            T[:, :, 0, k, k] = c
        return T, True
    if form == "rowdiv":
        T = np.zeros((E, Q, 3, 3, 1))
        for k in range(3):
            T[:, :, k, k, 0] = c
        return T, True
    Dr, Dc = (3 if rk == "gradient" else 1), (3 if ck == "gradient" else 1)
    shape = (E, Q, Cr) + ((Dr,) if rk == "gradient" else ()) + (Cc,) + ((Dc,) if ck == "gradient" else ())
    return rng.standard_normal(shape), True


@pytest.mark.parametrize("case", sorted(_CROSS_CASES))
def test_c_cross_kernel_matches_reference_cpp(ref, case):
    """General cross kernel and rectangular pattern against the reference's
    compiled CrossBilinearAssembler on identical inputs (bit-exact)."""
    p2, t2 = _small_p2(seed=7)
    X, W = rule4()
    (rq, Cr), (cq, Cc), rk, ck, _ = _CROSS_CASES[case]
    coef, tensor = _cross_inputs(case, p2, t2, X, W, seed=3)
    ip, ix, mine, Nr, Nc = pipeline.cross(p2, t2, X, W, (rq, Cr), (cq, Cc), rk, ck, coefficient=coef,
                                          tensor=tensor)
    (rs, rg, dx, *_), _ = pipeline.basis_geometry(p2, t2, X, W, rq)
    (cs, cg, *_), _ = pipeline.basis_geometry(p2, t2, X, W, cq)
    _, rd = pipeline.dofs_of(t2, p2.shape[1], 10 if rq else 4, Cr)
    _, cd = pipeline.dofs_of(t2, p2.shape[1], 10 if cq else 4, Cc)
    Cb = ref.CrossBilinearAssembler(rd, cd, rs, cs, dx, rg, cg)
    theirs, _ = Cb.assemble(np.ascontiguousarray(coef), rk, ck, 1)
    np.testing.assert_array_equal(ip, Cb.indptr)
    np.testing.assert_array_equal(ix, Cb.indices)
    assert (Cb.rows, Cb.columns) == (Nr, Nc)
    np.testing.assert_array_equal(mine, theirs)


def _linear_interpolant(p, t, quadratic, components, field):
    """Nodal interpolant of a vector-valued function (DOF order node*C + c)."""
    nodes = 10 if quadratic else 4
    used = np.unique(t[:nodes])
    vals = field(p[:, used])  # (C, nused)
    return np.ascontiguousarray(vals.T).ravel()


def test_cross_oracle_closed_forms():
    """Closed forms that pin orientation, components and directions of the
    cross kernel without the reference build:
    * Taylor-Hood B (test P1, trial vector P2, div(u) q): B u = int q div u,
      so u = (x,0,0) or (0,y,0) gives the P1 load vector sum_{e in a} |T_e|/4
      and u = (y,0,0) gives zero (a swapped direction/component fails);
    * the row-divergence block (test vector P2, trial P1) is B transposed;
    * P1 x P2 mass times the constant 1 is the same load vector;
    * P1 x P2 stiffness annihilates constants and x_P1^T K x_P2 = |Omega|."""
    p2, t2 = _small_p2(seed=11, n=2)
    X, W = rule4()
    E, Q = t2.shape[1], len(W)
    (rs, rg, dx, *_), _ = pipeline.basis_geometry(p2, t2, X, W, False)
    Nr, rd = pipeline.dofs_of(t2, p2.shape[1], 4, 1)
    load = np.zeros(Nr)
    np.add.at(load, rd[:, :, 0].ravel(), np.repeat(dx.sum(axis=1) / 4., 4))
    assert abs(load.sum() - 1.) < 1e-13
    T = np.zeros((E, Q, 1, 3, 3))
    for k in range(3):
        T[:, :, 0, k, k] = 1.
    ip, ix, vals, Nr, Nc = pipeline.cross(p2, t2, X, W, (False, 1), (True, 3), "value", "gradient",
                                          coefficient=T, tensor=True)
    from scipy.sparse import csr_matrix
    B = csr_matrix((vals, ix, ip), shape=(Nr, Nc))
    ux = _linear_interpolant(p2, t2, True, 3, lambda x: np.vstack([x[0], 0 * x[0], 0 * x[0]]))
    uy = _linear_interpolant(p2, t2, True, 3, lambda x: np.vstack([0 * x[0], x[1], 0 * x[0]]))
    uyx = _linear_interpolant(p2, t2, True, 3, lambda x: np.vstack([x[1], 0 * x[0], 0 * x[0]]))
    np.testing.assert_allclose(B @ ux, load, rtol=0, atol=1e-14)
    np.testing.assert_allclose(B @ uy, load, rtol=0, atol=1e-14)
    np.testing.assert_allclose(B @ uyx, 0., rtol=0, atol=1e-14)
    T2 = np.zeros((E, Q, 3, 3, 1))
    for k in range(3):
        T2[:, :, k, k, 0] = 1.
    ip2, ix2, vals2, Nr2, Nc2 = pipeline.cross(p2, t2, X, W, (True, 3), (False, 1), "gradient", "value",
                                               coefficient=T2, tensor=True)
    Bt = csr_matrix((vals2, ix2, ip2), shape=(Nr2, Nc2))
    assert (Nr2, Nc2) == (Nc, Nr)
    np.testing.assert_allclose(Bt.toarray(), B.T.toarray(), rtol=0, atol=1e-15)
    ip3, ix3, vals3, Nr3, Nc3 = pipeline.cross(p2, t2, X, W, (False, 1), (True, 1), "value", "value")
    M = csr_matrix((vals3, ix3, ip3), shape=(Nr3, Nc3))
    np.testing.assert_allclose(M @ np.ones(Nc3), load, rtol=0, atol=1e-15)
    ip4, ix4, vals4, Nr4, Nc4 = pipeline.cross(p2, t2, X, W, (False, 1), (True, 1), "gradient", "gradient")
    K = csr_matrix((vals4, ix4, ip4), shape=(Nr4, Nc4))
    np.testing.assert_allclose(K @ np.ones(Nc4), 0., rtol=0, atol=1e-13)
    x1 = _linear_interpolant(p2, t2, False, 1, lambda x: x[:1])
    x2 = _linear_interpolant(p2, t2, True, 1, lambda x: x[:1])
    assert abs(x1 @ (K @ x2) - 1.) < 1e-13


def test_composite_numbering_restatement():
    """The composite numbering (a literal restatement of the reference's node
    loop) equals its characterisation: DOFs ordered by (node, field,
    component), each field's components consecutive at a node."""
    p2, t2 = _small_p2(seed=2, n=2)
    conns = [t2[:10], t2[:4], t2[:4]]
    comps = [3, 1, 2]
    N, dofs = host.composite_dofs(p2.shape[1], conns, comps)
    keys = []
    for f, (c, C) in enumerate(zip(conns, comps)):
        for node in np.unique(c):
            for k in range(C):
                keys.append((int(node), f, k))
    assert N == len(keys)
    order = {key: i for i, key in enumerate(sorted(keys))}
    for f, (c, C) in enumerate(zip(conns, comps)):
        expect = np.array([[[order[(int(v), f, k)] for k in range(C)] for v in col] for col in c.T])
        np.testing.assert_array_equal(dofs[f], expect)


# ------------------------------------------------------------------ facets
def test_facet_topology_matches_restatement_and_invariants():
    """The vectorised mesh facet topology equals the oracle's loop
    restatement, facets are sorted unique triples in lexicographic order, and
    every element face maps to its facet."""
    from skfemntv.mesh import MeshTet
    for n in (1, 2, 3):
        ax = np.linspace(0, 1, n + 1)
        m = MeshTet.init_tensor(ax, ax, ax)
        facets, t2f, f2t = host.facet_topology(m.t)
        np.testing.assert_array_equal(m.facets, facets)
        np.testing.assert_array_equal(m.t2f, t2f)
        np.testing.assert_array_equal(m.f2t, f2t)
        assert np.all(np.diff(facets, axis=0) > 0)
        keys = [tuple(c) for c in facets.T]
        assert keys == sorted(keys) and len(set(keys)) == len(keys)
        for k, face in enumerate(host.TET_FACES):
            np.testing.assert_array_equal(np.sort(m.t[list(face)], axis=0), facets[:, t2f[k]])
        assert m.boundary_facets().size == 12 * n * n


@pytest.mark.parametrize("quadratic", [False, True])
def test_facet_oracle_closed_forms(quadratic):
    """Boundary of the jittered unit cube: sum dx = 6, int x.n ds = 3 |Omega|
    (divergence theorem), normals unit and outward; the load f = 1 gives
    sum_{facets of a} area/3 per P1 vertex, and on P2 zero at vertices and
    area/3 per facet at edge nodes."""
    p2, t2 = _small_p2(seed=3, n=2)
    p, t = (p2, t2) if quadratic else (p2[:, :27], t2[:4])
    fb = host.facet_basis(p, t, quadratic)
    dx, x, nrm = fb["dx"], fb["x"], fb["normals"]
    assert abs(dx.sum() - 6.) < 1e-13
    assert abs((dx * np.einsum("fqi,fqi->fq", x, nrm)).sum() - 3.) < 1e-13
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=-1), 1., rtol=0, atol=1e-15)
    F, Q, n = fb["shape"].shape
    grads = np.zeros((F, Q, n, 3))
    b = oracle.assemble_linear(fb["shape"], grads, dx, fb["dofs"], fb["N"], 1, value=np.array([1.]))
    area = dx.sum(axis=1)
    expect = np.zeros(fb["N"])
    par = fb["parents"]
    for f in range(F):
        corner = [t[c, par[f]] for c in host.TET_FACES[fb["local"][f]]]
        if quadratic:
            from oracle.host import TET_FACES  # noqa: F401
            edges = {(0, 1): 4, (1, 2): 5, (0, 2): 6, (0, 3): 7, (1, 3): 8, (2, 3): 9}
            lf = host.TET_FACES[fb["local"][f]]
            for i, j in ((lf[0], lf[1]), (lf[1], lf[2]), (lf[0], lf[2])):
                expect[fb["dofs"][f, edges[(min(i, j), max(i, j))], 0]] += area[f] / 3.
        else:
            for v in corner:
                expect[v] += area[f] / 3.
    np.testing.assert_allclose(b, expect, rtol=0, atol=1e-15)
