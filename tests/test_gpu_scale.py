"""GPU parity at the benchmark configurations' full sizes.

The small-mesh tests (test_gpu_assembly.py) pin every kernel and formulation;
these run the same checks where the launch layouts of the benchmarks apply:
one node row per lane, up to 64 rows per wave, 2-byte in-row offsets,
millions of hash-table edges and the XCD wave placement.

* C1 P1 Laplace + load, 32^3: pattern bit-exact, every value <= 1e-12
  against the oracle (bilinear_form_assembler.cpp:39-140,
  linear_form_assembler.cpp:41-137), FP32 products <= 1e-6.
* C2 P2 anisotropic diffusion, 64^3 jittered: the P2 mesh numbering, the
  pattern and every FP64 value against the oracle (cross_bilinear_assembler.
  cpp:19-107,207-341), FP32 products <= 1e-6.
* C3 vector-P2 elasticity, 48^3 jittered: the full DOF-level pattern (the
  node pattern from the oracle, expanded by the interleaved numbering of
  basis.py:1768), FP64 and FP32 values of >= 1e5 DOF rows recomputed by the
  oracle from every element that touches them.
* C4 N1 curl-curl + kappa mass, 96^3: edge ids, element edges and signs
  bit-exact (edge_topology.cpp:30-84), the edge pattern bit-exact
  (edge_topology.cpp:86-154), FP64 values of >= 1e5 rows recomputed by the
  oracle (integrate_tet_n1, edge_topology.cpp:156-233, and the element-major
  scatter of hcurl_assembler.py:94-97).

A row of the sampled comparisons is complete because the oracle assembles
every element incident to it, in ascending element order.
"""

import sys

import numpy as np
import pytest

import oracle
from oracle import host, pipeline

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

FP64_TOL = 1e-12
FP32_TOL = 1e-6


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def kuhn(n, jitter=0.0, seed=0):
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    if jitter:
        p = host.jitter_interior(p, 1. / n, jitter, seed)
    return p, t


def spd_tensors(E, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((E, 3, 3))
    q, r = np.linalg.qr(g)
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    lam = 10. ** rng.uniform(-2, 2, size=(E, 3))
    return np.einsum("eij,ej,ekj->eik", q, lam, q)


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.))


def rows_of(ip, rows):
    """Flat CSR positions of the given rows (ascending)."""
    lo, hi = ip[rows], ip[rows + 1]
    lens = hi - lo
    start = np.repeat(lo - np.concatenate(([0], np.cumsum(lens)[:-1])), lens)
    return start + np.arange(int(lens.sum()))


def compare_rows(K, ip, ix, vals, rows, tol):
    """Columns bit-exact and values within tol on the given rows of K
    against an oracle CSR (ip, ix, vals) whose rows are complete there."""
    Kip = np.asarray(K.indptr, dtype=np.int64)
    np.testing.assert_array_equal(np.diff(Kip)[rows], np.diff(ip)[rows])
    pk, po = rows_of(Kip, rows), rows_of(ip, rows)
    np.testing.assert_array_equal(np.asarray(K.indices)[pk].astype(np.int64), ix[po])
    return rel(K.data[pk], vals[po])


def test_c1_p1_laplace_load_32_full():
    import skfemntv as sk
    from skfemntv._assembly import NativeBilinearForm
    from skfemntv.helpers import ddot, dot, grad
    say("C1 32^3 P1 laplace + load: full pattern and values")
    p, t = kuhn(32)
    ax = np.linspace(0., 1., 33)
    mesh = sk.MeshTet.init_tensor(ax, ax, ax)
    np.testing.assert_array_equal(mesh.t, t)
    np.testing.assert_array_equal(mesh.p, p)
    basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1))
    K = sk.asm(sk.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v))), basis)
    b = sk.asm(sk.LinearForm(lambda v, w: dot(w.source, v)), basis, source=np.array([1.]))
    X, W = host.tet_rule_4()
    E = t.shape[1]
    ip, ix, vals, N = pipeline.bilinear(p, t, X, W, False, gradient=np.ones((E, 4)), threads=8)
    assert len(ix) == 513313
    np.testing.assert_array_equal(np.asarray(K.indptr, dtype=np.int64), ip)
    np.testing.assert_array_equal(np.asarray(K.indices, dtype=np.int64), ix)
    e64 = rel(K.data, vals)
    assert e64 <= FP64_TOL
    eb = rel(b, pipeline.load_vector(p, t, X, W, False, [1.]))
    assert eb <= FP64_TOL
    K32 = NativeBilinearForm(basis).assemble(gradient=np.ones((E, 1)), precision="fp32")
    e32 = rel(K32.data, vals)
    say(f"C1 fp64 {e64:.2e} load {eb:.2e} fp32 {e32:.2e}")
    assert e32 <= FP32_TOL


def test_c2_p2_anisotropic_64_full():
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm
    from skfemntv.helpers import dot, grad, mul
    say("C2 64^3 P2 anisotropic: mesh numbering, full pattern and values")
    p, t = kuhn(64, 0.2, 0)
    p2, t2 = host.p2_mesh(p, t)
    mesh = sk.MeshTet2.from_mesh(sk.MeshTet(p, t))
    np.testing.assert_array_equal(mesh.t, t2)
    np.testing.assert_array_equal(mesh.p, p2)
    X, W = host.tet_rule_4()
    basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=(X, W))
    E = t.shape[1]
    A = spd_tensors(E, 1)
    A_eq = np.ascontiguousarray(np.broadcast_to(A[:, None], (E, 4, 3, 3)))

    @sk.BilinearForm
    def aniso(u, v, w):
        return dot(mul(w.diffusion, grad(u)), grad(v))

    K = sk.asm(aniso, basis, diffusion=np.moveaxis(A_eq, (2, 3), (0, 1)))
    ip, ix, vals, N = pipeline.tensor(p2, t2, X, W, True, A_eq, threads=8)
    assert len(ix) == 60859905
    np.testing.assert_array_equal(np.asarray(K.indptr, dtype=np.int64), ip)
    np.testing.assert_array_equal(np.asarray(K.indices, dtype=np.int64), ix)
    e64 = rel(K.data, vals)
    assert e64 <= FP64_TOL
    form = NativeCrossBilinearForm(basis, basis)
    K32 = form.assemble_tensor("gradient", "gradient", A_eq[:, :, None, :, None, :], precision="fp32")
    e32 = rel(K32.data, vals)
    say(f"C2 fp64 {e64:.2e} fp32 {e32:.2e}")
    assert e32 <= FP32_TOL


def test_c3_vector_p2_elasticity_48_pattern_and_rows():
    import skfemntv as sk
    from skfemntv._assembly import NativeBilinearForm
    from skfemntv.helpers import ddot, div, sym_grad
    say("C3 48^3 vector P2 elasticity: full pattern, sampled rows")
    p, t = kuhn(48, 0.15, 0)
    p2, t2 = host.p2_mesh(p, t)
    mesh = sk.MeshTet2.from_mesh(sk.MeshTet(p, t))
    np.testing.assert_array_equal(mesh.t, t2)
    X, W = host.tet_rule_4()
    basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=3), quadrature=(X, W))
    E, Q = t.shape[1], 4
    rng = np.random.default_rng(1)
    Ey = np.exp(rng.standard_normal(E))
    nu = rng.uniform(0.2, 0.45, E)
    lam = Ey * nu / ((1 + nu) * (1 - 2 * nu))
    mu = Ey / (2 * (1 + nu))

    @sk.BilinearForm
    def elasticity(u, v, w):
        return 2. * w.mu * ddot(sym_grad(u), sym_grad(v)) + w.lam * (div(u) * div(v))

    K = sk.asm(elasticity, basis, mu=mu[:, None], lam=lam[:, None])
    nn = p2.shape[1]
    N = 3 * nn
    assert K.shape == (N, N) and K.nnz == 231797385
    # node pattern from the oracle, expanded: DOF row 3r+c holds columns
    # 3j+d for the row's column nodes j in order, d = 0..2
    ipn, ixn, _ = oracle.pattern(np.ascontiguousarray(t2.T), nrows=nn)
    lens = np.diff(ipn)
    Kip = np.asarray(K.indptr, dtype=np.int64)
    np.testing.assert_array_equal(np.diff(Kip), np.repeat(3 * lens, 3))
    Kix = np.asarray(K.indices)
    for r0 in range(0, nn, 200000):
        r1 = min(nn, r0 + 200000)
        k0, k1 = ipn[r0], ipn[r1]
        rowk = np.repeat(np.arange(r0, r1), lens[r0:r1])
        j = np.arange(k0, k1) - ipn[rowk]
        for c in range(3):
            base = 9 * ipn[rowk] + c * 3 * lens[rowk] + 3 * j
            for d in range(3):
                np.testing.assert_array_equal(Kix[base + d].astype(np.int64), 3 * ixn[k0:k1] + d)
    del Kix
    # >= 1e5 DOF rows: every 25th node, all its components
    nodes = np.arange(0, nn, 25)
    rows = (3 * nodes[:, None] + np.arange(3)).ravel()
    assert rows.size >= 100000
    touch = np.zeros(nn, dtype=bool)
    touch[nodes] = True
    es = np.flatnonzero(touch[t2].any(axis=0))
    ts = np.ascontiguousarray(t2[:, es])
    (shape, grads, dx, *_), _ = pipeline.basis_geometry(p2, ts, X, W, True)
    _, dofs_all = host.element_dofs(t2, nn, 3)
    dofs = np.ascontiguousarray(dofs_all[:, es].T)
    ips, ixs, sc = oracle.pattern(dofs, nrows=N)
    coef = dict(symgrad=np.repeat(2 * mu[es, None], Q, 1), divergence=np.repeat(lam[es, None], Q, 1))
    vals = oracle.assemble_bilinear(shape, grads, dx, sc, len(ixs), 3, **coef)
    e64 = compare_rows(K, ips, ixs, vals, rows, FP64_TOL)
    assert e64 <= FP64_TOL
    K32 = NativeBilinearForm(basis).assemble(symmetric_gradient=2 * mu[:, None], divergence=lam[:, None],
                                             precision="fp32")
    e32 = compare_rows(K32, ips, ixs, vals, rows, FP32_TOL)
    say(f"C3 sampled rows {rows.size}: fp64 {e64:.2e} fp32 {e32:.2e}")
    assert e32 <= FP32_TOL
    # the 16-row layout of the vector-P2 MFMA gather at full size
    form = NativeBilinearForm(basis)
    form.plan.rebuild(32768, row_lanes=4)
    form.plan._tuned = True
    for prec, tol in (("fp64", FP64_TOL), ("fp32", FP32_TOL)):
        Km = form.assemble(symmetric_gradient=2 * mu[:, None], divergence=lam[:, None], precision=prec)
        em = compare_rows(Km, ips, ixs, vals, rows, tol)
        say(f"C3 MFMA layout ({form.plan.waves} waves) {prec}: {em:.2e}")
        assert em <= tol


def test_c4_n1_96_edges_pattern_and_rows():
    import skfemntv as sk
    say("C4 96^3 N1: edges and signs, full pattern, sampled rows")
    p, t = kuhn(96)
    mesh = sk.MeshTet(p, t)
    basis = sk.AffineTetN1Basis(mesh)
    edges, ee, signs = oracle.edge_topology(t[:4])
    assert edges.shape[1] == 6276384
    topo = basis.dof_map.topology
    np.testing.assert_array_equal(topo.edges, edges)
    np.testing.assert_array_equal(topo.element_edges, ee)
    np.testing.assert_array_equal(basis.basis_signs, signs)
    E, nE = t.shape[1], edges.shape[1]
    kappa = np.random.default_rng(2).uniform(0.1, 10., E)
    K = sk.TetN1Assembler(basis).assemble_maxwell(mass_coefficient=kappa[:, None], curl_coefficient=1.)
    ip, ix, sc = oracle.pattern(np.ascontiguousarray(ee.T), nrows=nE)
    assert len(ix) == 102159648
    np.testing.assert_array_equal(np.asarray(K.indptr, dtype=np.int64), ip)
    np.testing.assert_array_equal(np.asarray(K.indices, dtype=np.int64), ix)
    # every 50th edge row, from the elements that touch them
    rows = np.arange(0, nE, 50)
    assert rows.size >= 100000
    touch = np.zeros(nE, dtype=bool)
    touch[rows] = True
    es = np.flatnonzero(touch[ee].any(axis=0))
    X, W = host.tet_rule_duffy(2)
    el = np.empty((es.size, 6, 6))
    for c0 in range(0, es.size, 100000):
        sl = es[c0:c0 + 100000]
        dx, values, curls, _ = host.n1_basis(p, t[:, sl], X, W, signs[:, sl])
        Q = dx.shape[1]
        el[c0:c0 + sl.size] = oracle.integrate_tet_n1(values, curls, dx, np.repeat(kappa[sl, None], Q, 1),
                                                      np.ones((sl.size, Q)))
    data = oracle.scatter_add(sc[es], el, len(ix))
    e64 = compare_rows(K, ip, ix, data, rows, FP64_TOL)
    say(f"C4 sampled rows {rows.size}: fp64 {e64:.2e}")
    assert e64 <= FP64_TOL


@pytest.mark.parametrize("case", ["p1", "p2_tensor", "vp2", "n1"])
def test_one_row_per_lane_layout_on_small_meshes(case):
    """min_waves = 0 keeps one row per lane on small meshes, so the waves
    hold up to 64 rows (R > 1, S = 1), the layout of the benchmark sizes;
    4-node scalar forms then take the element-table gather.  Values must
    match the oracle and be bit-identical to the default small-mesh layout
    (rows split over lanes)."""
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, Plan, upload
    from skfemntv.helpers import ddot, div, grad, sym_grad
    say(f"one row per lane layout: {case}")
    X, W = host.tet_rule_4()
    if case == "n1":
        p, t = kuhn(5, 0.1, 31)
        basis = sk.AffineTetN1Basis(sk.MeshTet(p, t))
        asm_ = sk.TetN1Assembler(basis)
        E = t.shape[1]
        kappa = np.random.default_rng(3).uniform(0.1, 10., E)
        ref = asm_.assemble_maxwell(mass_coefficient=kappa[:, None], curl_coefficient=1.).data.copy()
        asm_.plan = Plan(basis._edges_dev, E, 6, basis.N, 1, min_waves=0)
        assert asm_.plan.waves < basis.N
        K = asm_.assemble_maxwell(mass_coefficient=kappa[:, None], curl_coefficient=1.)
        ip, ix, data, *_ = pipeline.n1_maxwell(p, t, np.repeat(kappa[:, None], len(basis.W), 1),
                                              np.ones((E, len(basis.W))))
        np.testing.assert_array_equal(np.asarray(K.indices, dtype=np.int64), ix)
        assert rel(K.data, data) <= FP64_TOL
        np.testing.assert_array_equal(K.data, ref)
        return
    if case == "p1":
        p, t = kuhn(8, 0.2, 32)
        basis = sk.Basis(sk.MeshTet(p, t), sk.ElementVector(sk.ElementTetP1(), dim=1))
        E = t.shape[1]
        ip, ix, vals, N = pipeline.bilinear(p, t, X, W, False, gradient=np.ones((E, 4)))
        # a new form object per call: asm caches its assembler (and plan) per form
        run = lambda: sk.asm(sk.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v))), basis).data.copy()  # noqa: E731
        n, C = 4, 1
    elif case == "p2_tensor":
        p, t = kuhn(5, 0.2, 33)
        p, t = host.p2_mesh(p, t)
        basis = sk.Basis(sk.MeshTet2(p, t), sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=(X, W))
        E = t.shape[1]
        A = spd_tensors(E, 7)
        A_eq = np.ascontiguousarray(np.broadcast_to(A[:, None], (E, 4, 3, 3)))
        ip, ix, vals, N = pipeline.tensor(p, t, X, W, True, A_eq)
        cross = NativeCrossBilinearForm(basis, basis)
        Ad = upload(A)
        run = lambda: cross.assemble_tensor_device(Ad, symmetric=True)[:len(vals)].cpu().numpy()  # noqa: E731
        n, C = 10, 1
    else:
        p, t = kuhn(4, 0.15, 34)
        p, t = host.p2_mesh(p, t)
        basis = sk.Basis(sk.MeshTet2(p, t), sk.ElementVector(sk.ElementTetP2(), dim=3), quadrature=(X, W))
        E = t.shape[1]
        rng = np.random.default_rng(5)
        mu, lam = np.exp(rng.standard_normal(E)), np.exp(rng.standard_normal(E))
        def form():
            return sk.BilinearForm(lambda u, v, w: 2. * w.mu * ddot(sym_grad(u), sym_grad(v))
                                   + w.lam * (div(u) * div(v)))
        ip, ix, vals, N = pipeline.bilinear(p, t, X, W, True, components=3, symgrad=np.repeat(2 * mu[:, None], 4, 1),
                                            divergence=np.repeat(lam[:, None], 4, 1))
        run = lambda: sk.asm(form(), basis, mu=mu[:, None], lam=lam[:, None]).data.copy()  # noqa: E731
        n, C = 10, 3
    default = run()
    d = basis.device_data()
    plan = Plan(d["node_conn"], E, n, d["nnodes"], C, lds_budget=basis._plan.lds_budget, min_waves=0)
    assert basis._plan.waves > plan.waves
    if case == "p2_tensor":
        cross.plan = plan
    else:
        basis._plan = plan
    got = run()
    assert rel(got, vals) <= FP64_TOL
    if case == "p1":
        # the element-table formulation (automatic at one row per lane for
        # 4-node elements) rounds differently from the quadrature gather
        assert rel(got, default) <= FP64_TOL
    else:
        np.testing.assert_array_equal(got, default)
