"""GPU parity of the HIP assembly path against the CPU oracle.

Integer outputs (CSR pattern, edge ids, signs) must be bit-exact; FP64
values within 1e-12 relative Frobenius; the FP32-products form within 1e-6.
Each case prints its name before touching the GPU.
"""

import sys

import numpy as np
import pytest

import oracle
from oracle import host, pipeline

pytestmark = pytest.mark.gpu

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


def check_csr(K, ip, ix, vals, tol=FP64_TOL):
    K = K.tocsr()
    np.testing.assert_array_equal(np.asarray(K.indptr, dtype=np.int64), ip)
    np.testing.assert_array_equal(np.asarray(K.indices, dtype=np.int64), ix)
    err = pipeline.relative_frobenius(K.data, vals)
    assert err <= tol, err
    return err


def test_p1_laplace_and_load_match_oracle():
    import skfemntv as sk
    from skfemntv.helpers import ddot, dot, grad
    for n, jit in ((3, 0.0), (4, 0.2)):
        say(f"p1 laplace n={n} jitter={jit}")
        p, t = kuhn(n, jit, seed=n)
        mesh = sk.MeshTet(p, t)
        basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1))

        @sk.BilinearForm
        def laplace(u, v, w):
            return ddot(grad(u), grad(v))

        @sk.LinearForm
        def load(v, w):
            return dot(w.source, v)

        K = sk.asm(laplace, basis)
        b = sk.asm(load, basis, source=np.array([1.]))
        X, W = host.tet_rule_4()
        E, Q = t.shape[1], 4
        ip, ix, vals, N = pipeline.bilinear(p, t, X, W, False, gradient=np.ones((E, Q)))
        check_csr(K, ip, ix, vals)
        ref_b = pipeline.load_vector(p, t, X, W, False, [1.])
        assert pipeline.relative_frobenius(b, ref_b) <= FP64_TOL
        # closed forms: constants in the kernel, sum of load = volume
        assert np.abs(K @ np.ones(N)).max() <= 1e-12 * np.abs(K.data).max()
        assert abs(b.sum() - 1.) <= 1e-13


def test_p1_laplace_dot_spelling_is_accepted():
    import skfemntv as sk
    from skfemntv.helpers import ddot, dot, grad
    say("p1 dot(grad,grad) spelling")
    p, t = kuhn(2)
    basis = sk.Basis(sk.MeshTet(p, t), sk.ElementVector(sk.ElementTetP1(), dim=1))
    K1 = sk.asm(sk.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v))), basis).copy()
    K2 = sk.asm(sk.BilinearForm(lambda u, v, w: dot(grad(u), grad(v))), basis)
    assert (K1 - K2).nnz == 0 or np.abs((K1 - K2).data).max() == 0.


@pytest.mark.parametrize("layout", ["recommended", "native"])
def test_p2_anisotropic_fp64_matches_oracle(layout):
    import skfemntv as sk
    from skfemntv.helpers import dot, grad, mul
    say(f"p2 anisotropic layout={layout}")
    n = 4
    p, t = kuhn(n, 0.2, seed=7)
    p2, t2 = host.p2_mesh(p, t)
    mesh = sk.MeshTet2(p2, t2)
    X, W = host.tet_rule_4()
    basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=(X, W))
    E, Q = t.shape[1], 4
    A = spd_tensors(E, 11)
    A_eq = np.broadcast_to(A[:, None], (E, Q, 3, 3))
    raw = np.moveaxis(A_eq, (2, 3), (0, 1)) if layout == "recommended" else A_eq

    @sk.BilinearForm
    def aniso(u, v, w):
        return dot(mul(w.diffusion, grad(u)), grad(v))

    K = sk.asm(aniso, basis, diffusion=raw)
    ip, ix, vals, N = pipeline.tensor(p2, t2, X, W, True, np.ascontiguousarray(A_eq))
    err = check_csr(K, ip, ix, vals)
    say(f"p2 anisotropic fp64 rel err {err:.3e}")
    # the tensor path returns a fresh matrix: a second call leaves the first intact
    first = K.data.copy()
    K2 = sk.asm(aniso, basis, diffusion=2. * raw)
    np.testing.assert_array_equal(K.data, first)
    assert K2.data is not K.data
    np.testing.assert_allclose(K2.data, 2. * first, rtol=1e-14, atol=0.)


def test_p2_anisotropic_fp32_products_within_1e6():
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, upload
    say("p2 anisotropic fp32 products")
    p, t = kuhn(4, 0.2, seed=3)
    p2, t2 = host.p2_mesh(p, t)
    X, W = host.tet_rule_4()
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=(X, W))
    E = t.shape[1]
    A = spd_tensors(E, 5)
    form = NativeCrossBilinearForm(basis, basis)
    v64 = form.assemble_tensor_device(upload(A), precision="fp64").clone().cpu().numpy()
    v32 = form.assemble_tensor_device(upload(A), precision="fp32").clone().cpu().numpy()
    ip, ix, vals, N = pipeline.tensor(p2, t2, X, W, True, np.ascontiguousarray(np.broadcast_to(A[:, None], (E, 4, 3, 3))))
    e64 = pipeline.relative_frobenius(v64[:len(vals)], vals)
    e32 = pipeline.relative_frobenius(v32[:len(vals)], vals)
    say(f"fp64 {e64:.3e} fp32 {e32:.3e}")
    assert e64 <= FP64_TOL
    assert e32 <= FP32_TOL


@pytest.mark.parametrize("quadratic", [False, True])
def test_symmetric_tensor_formulations_match_oracle(quadratic):
    """Element-table gather and quadrature gather, FP64 and FP32 products,
    on a symmetric per-element tensor (the benchmark's form)."""
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, upload
    say(f"symmetric tensor formulations quadratic={quadratic}")
    p, t = kuhn(4, 0.2, seed=13)
    X, W = host.tet_rule_4()
    if quadratic:
        p, t = host.p2_mesh(p, t)
        basis = sk.Basis(sk.MeshTet2(p, t), sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=(X, W))
    else:
        basis = sk.Basis(sk.MeshTet(p, t), sk.ElementVector(sk.ElementTetP1(), dim=1), quadrature=(X, W))
    E = t.shape[1]
    A = spd_tensors(E, 17)
    A = 0.5 * (A + np.swapaxes(A, 1, 2))
    ip, ix, vals, N = pipeline.tensor(p, t, X, W, quadratic,
                                      np.ascontiguousarray(np.broadcast_to(A[:, None], (E, 4, 3, 3))))
    form = NativeCrossBilinearForm(basis, basis)
    Ad = upload(A)
    out, raw = {}, {}
    # (element table, mfma, row split, row form): row form 1 (barycentric
    # rows, P2 with the 4-point rule) and split 3 (concurrent parts) round
    # differently; the others must agree bit for bit across splits
    for table, mfma, split, form_ in ((1, 0, 0, 0), (0, 0, 0, 0), (0, 0, 1, 0), (0, 0, 2, 0), (0, 1, 0, 0),
                                      (0, 2, 0, 0), (0, 0, 0, 1), (0, 0, 3, 1), (0, 0, 3, 0)):
        form.plan.set_tuning(element_table=table, mfma=mfma, row_split=split, row_form=form_)
        for prec in ("fp64", "fp32"):
            v = form.assemble_tensor_device(Ad, precision=prec, symmetric=True).clone().cpu().numpy()[:len(vals)]
            out[(table, mfma, split, form_, prec)] = pipeline.relative_frobenius(v, vals)
            raw[(table, mfma, split, form_, prec)] = v
    # the row splits change only which lane computes or adds an entry: same bits
    for prec in ("fp64", "fp32"):
        np.testing.assert_array_equal(raw[(0, 0, 1, 0, prec)], raw[(0, 0, 0, 0, prec)])
        np.testing.assert_array_equal(raw[(0, 0, 2, 0, prec)], raw[(0, 0, 0, 0, prec)])
    form.plan.set_tuning(element_table=-1, mfma=-1, row_split=0, row_form=-1)
    say("rel errors " + ", ".join(f"table={k[0]} mfma={k[1]} split={k[2]} form={k[3]} {k[4]}: {e:.2e}"
                                  for k, e in out.items()))
    for key, e in out.items():
        assert e <= (FP64_TOL if key[-1] == "fp64" else FP32_TOL), (key, e)


@pytest.mark.parametrize("components", [1, 3])
def test_row_splits_agree_at_small_lds_budgets(components):
    """Small span budgets put several lanes on one node row (S = 2..16 for
    scalar P2, S = 2..4 for 3-component fields): the incidence, entry and
    column splits give the same bits, and match the oracle."""
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, Plan, upload
    from skfemntv.helpers import ddot, div, sym_grad
    say(f"row splits at small LDS budgets, components={components}")
    p, t = kuhn(3, 0.15, seed=21)
    p2, t2 = host.p2_mesh(p, t)
    X, W = host.tet_rule_4()
    E, Q = t.shape[1], 4
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=components), quadrature=(X, W))
    d = basis.device_data()
    if components == 1:
        A = spd_tensors(E, 5)
        A = 0.5 * (A + np.swapaxes(A, 1, 2))
        ip, ix, vals, N = pipeline.tensor(p2, t2, X, W, True,
                                          np.ascontiguousarray(np.broadcast_to(A[:, None], (E, Q, 3, 3))))
        form = NativeCrossBilinearForm(basis, basis)
        Ad = upload(A)
        budgets = (2048, 4096)
    else:
        rng = np.random.default_rng(3)
        mu, lam = np.exp(rng.standard_normal(E)), np.exp(rng.standard_normal(E))

        def elasticity(u, v, w):
            return 2. * w.mu * ddot(sym_grad(u), sym_grad(v)) + w.lam * (div(u) * div(v))

        ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, components=3,
                                            symgrad=np.repeat(2 * mu[:, None], Q, 1),
                                            divergence=np.repeat(lam[:, None], Q, 1))
        budgets = (2048, 8192)
    for budget in budgets:
        plan = Plan(d["node_conn"], E, 10, d["nnodes"], components, lds_budget=budget)
        raw = {}
        for split in (0, 1, 2):
            plan.set_tuning(row_split=split, row_form=0)
            if components == 1:
                form.plan = plan
                v = form.assemble_tensor_device(Ad, precision="fp64", symmetric=True).clone().cpu().numpy()
                raw[split] = v[:len(vals)]
                assert pipeline.relative_frobenius(raw[split], vals) <= FP64_TOL
            else:
                basis._plan = plan
                # a new form object: asm caches its assembler (and plan) per form
                K = sk.asm(sk.BilinearForm(elasticity), basis, mu=mu[:, None], lam=lam[:, None])
                assert K.nnz == plan.nnz
                check_csr(K, ip, ix, vals)
                raw[split] = K.data.copy()
        say(f"budget {budget}: waves {plan.waves}")
        np.testing.assert_array_equal(raw[1], raw[0])
        np.testing.assert_array_equal(raw[2], raw[0])
        if components == 1:
            # barycentric rows with concurrent parts: within the FP64 bound
            plan.set_tuning(row_split=3, row_form=1)
            form.plan = plan
            v = form.assemble_tensor_device(Ad, precision="fp64", symmetric=True).clone().cpu().numpy()
            assert pipeline.relative_frobenius(v[:len(vals)], vals) <= FP64_TOL


def test_vector_p2_elasticity_matches_oracle():
    import skfemntv as sk
    from skfemntv.helpers import ddot, div, dot, grad, sym_grad
    say("vector p2 elasticity")
    p, t = kuhn(3, 0.15, seed=2)
    p2, t2 = host.p2_mesh(p, t)
    X, W = host.tet_rule_4()
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=3), quadrature=(X, W))
    E, Q = t.shape[1], 4
    rng = np.random.default_rng(9)
    Ey = np.exp(rng.standard_normal(E))
    nu = rng.uniform(0.2, 0.45, E)
    lam = Ey * nu / ((1 + nu) * (1 - 2 * nu))
    mu = Ey / (2 * (1 + nu))

    @sk.BilinearForm
    def elasticity(u, v, w):
        return 2. * w.mu * ddot(sym_grad(u), sym_grad(v)) + w.lam * (div(u) * div(v))

    K = sk.asm(elasticity, basis, mu=mu[:, None], lam=lam[:, None])
    ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, components=3,
                                        symgrad=np.repeat(2 * mu[:, None], Q, 1),
                                        divergence=np.repeat(lam[:, None], Q, 1))
    err = check_csr(K, ip, ix, vals)
    say(f"elasticity rel err {err:.3e}")
    # rigid translations are in the kernel
    for c in range(3):
        r = np.zeros(N)
        r[c::3] = 1.
        assert np.abs(K @ r).max() <= 1e-11 * np.abs(K.data).max()
    # with a mass term (rho u . v), a gradient term and FP32 products
    rho = np.exp(rng.standard_normal(E))
    kap = np.exp(rng.standard_normal(E))

    @sk.BilinearForm
    def dynamic(u, v, w):
        return (2. * w.mu * ddot(sym_grad(u), sym_grad(v)) + w.lam * (div(u) * div(v))
                + w.kap * ddot(grad(u), grad(v)) + w.rho * dot(u, v))

    say("vector p2 elasticity + gradient + mass")
    K2 = sk.asm(dynamic, basis, mu=mu[:, None], lam=lam[:, None], kap=kap[:, None], rho=rho[:, None])
    ip, ix, vals2, N = pipeline.bilinear(p2, t2, X, W, True, components=3,
                                         value=np.repeat(rho[:, None], Q, 1),
                                         gradient=np.repeat(kap[:, None], Q, 1),
                                         symgrad=np.repeat(2 * mu[:, None], Q, 1),
                                         divergence=np.repeat(lam[:, None], Q, 1))
    err2 = check_csr(K2, ip, ix, vals2)
    say(f"elasticity + mass rel err {err2:.3e}")


def test_vector_p2_block_forms_match_oracle():
    """Vector-P2 gather, both 3x3 block forms (row form 0: 4-point sums;
    1: barycentric pair table), incidence and concurrent splits, at a small
    span budget (two and four lanes per node row) and the default one,
    FP64 and FP32 products, with value, gradient, symmetric-gradient and
    divergence terms: every combination against the oracle
    (bilinear_form_assembler.cpp:79-105)."""
    import skfemntv as sk
    from skfemntv._assembly import NativeBilinearForm, Plan, upload
    say("vector p2 block forms x splits x budgets")
    p, t = kuhn(3, 0.15, seed=31)
    p2, t2 = host.p2_mesh(p, t)
    X, W = host.tet_rule_4()
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=3), quadrature=(X, W))
    E, Q = t.shape[1], 4
    rng = np.random.default_rng(32)
    coef = {k: np.exp(rng.standard_normal(E)) for k in ("value", "gradient", "symmetric_gradient", "divergence")}
    oracle_kw = {"value": "value", "gradient": "gradient", "symmetric_gradient": "symgrad", "divergence": "divergence"}
    ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, components=3,
                                        **{oracle_kw[k]: np.repeat(v[:, None], Q, 1) for k, v in coef.items()})
    dev = {k: upload(v) for k, v in coef.items()}
    d = basis.device_data()
    assert d["rtab_barycentric"]
    form = NativeBilinearForm(basis)
    results = {}
    for budget in (8192, None):
        plan = Plan(d["node_conn"], E, 10, d["nnodes"], 3, lds_budget=budget) if budget else form.plan
        form.plan = plan
        for row_form in (0, 1):
            for split in (0, 3):
                plan.set_tuning(row_form=row_form, row_split=split)
                for prec in ("fp64", "fp32"):
                    v = form.assemble_device(precision=prec, **dev)[:len(vals)].cpu().numpy()
                    err = pipeline.relative_frobenius(v, vals)
                    results[(budget, row_form, split, prec)] = err
                    assert err <= (FP64_TOL if prec == "fp64" else FP32_TOL), (budget, row_form, split, prec, err)
    say("rel errors " + ", ".join(f"{k}: {e:.1e}" for k, e in results.items()))


def test_vector_p2_mfma_layout_matches_oracle():
    """The vector-P2 MFMA gather (plan layout with 4 lanes per node row):
    16-row waves (min_waves 0), the split layouts of a small mesh (fewer rows
    per wave) and a small span budget, FP64 (f64 MFMA) and FP32 products
    (f32 MFMA, FP32 accumulation), with value, gradient, symmetric-gradient
    and divergence terms, and with one kind at a time; plus the fallback of
    a per-point coefficient to the default layout
    (bilinear_form_assembler.cpp:79-105)."""
    import skfemntv as sk
    from skfemntv._assembly import NativeBilinearForm, Plan, upload
    say("vector p2 MFMA layout")
    p, t = kuhn(4, 0.15, seed=41)
    p2, t2 = host.p2_mesh(p, t)
    X, W = host.tet_rule_4()
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=3), quadrature=(X, W))
    E, Q = t.shape[1], 4
    rng = np.random.default_rng(42)
    coef = {k: np.exp(rng.standard_normal(E)) for k in ("value", "gradient", "symmetric_gradient", "divergence")}
    oracle_kw = {"value": "value", "gradient": "gradient", "symmetric_gradient": "symgrad", "divergence": "divergence"}
    d = basis.device_data()
    form = NativeBilinearForm(basis)
    cases = [tuple(coef)] + [(k,) for k in ("value", "symmetric_gradient", "divergence")]
    results = {}
    for kinds in cases:
        ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, components=3,
                                            **{oracle_kw[k]: np.repeat(coef[k][:, None], Q, 1) for k in kinds})
        dev = {k: upload(coef[k]) for k in kinds}
        for budget, min_waves in ((32768, 0), (32768, 2048), (8192, 0)):
            plan = Plan(d["node_conn"], E, 10, d["nnodes"], 3, lds_budget=budget, min_waves=min_waves, row_lanes=4)
            plan._tuned = True
            form.plan = plan
            for prec in ("fp64", "fp32"):
                v = form.assemble_device(precision=prec, **dev)[:len(vals)].cpu().numpy()
                err = pipeline.relative_frobenius(v, vals)
                results[(kinds if len(kinds) == 1 else "all", budget, min_waves, prec)] = err
                assert err <= (FP64_TOL if prec == "fp64" else FP32_TOL), (kinds, budget, min_waves, prec, err)
            assert plan.row_lanes == 4
    say("rel errors " + ", ".join(f"{k}: {e:.1e}" for k, e in results.items()))
    # a per-point coefficient is outside the MFMA layout: the plan returns to
    # the default layout and the result still matches the oracle
    sg = np.exp(rng.standard_normal((E, Q)))
    ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, components=3, symgrad=sg)
    v = form.assemble_device(symmetric_gradient=upload(sg), per_qp=True)[:len(vals)].cpu().numpy()
    assert form.plan.row_lanes == 0
    assert pipeline.relative_frobenius(v, vals) <= FP64_TOL


def test_p2_mass_and_perqp_tensor_match_oracle():
    import skfemntv as sk
    from skfemntv.helpers import dot, grad, mul
    say("p2 mass + per-qp nonsymmetric tensor")
    p, t = kuhn(3, 0.2, seed=4)
    p2, t2 = host.p2_mesh(p, t)
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=1))  # default Keast rule
    X, W = host.tet_rule_keast11()
    E, Q = t.shape[1], 11

    @sk.BilinearForm
    def mass(u, v, w):
        return 0.75 * dot(u, v)

    M = sk.asm(mass, basis)
    ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, value=np.full((E, Q), 0.75))
    check_csr(M, ip, ix, vals)
    x = np.moveaxis(basis.global_coordinates, -1, 0)
    T = np.empty((3, 3, E, Q))
    T[0, 0] = 1. + x[0]; T[0, 1] = 0.3 + 0.1 * x[1]; T[0, 2] = 0.
    T[1, 0] = -0.2; T[1, 1] = 2. - 0.4 * x[0]; T[1, 2] = 0.1 * x[2]
    T[2, 0] = 0.05; T[2, 1] = 0.; T[2, 2] = 1.5

    @sk.BilinearForm
    def aniso(u, v, w):
        return dot(mul(w.diffusion, grad(u)), grad(v))

    K = sk.asm(aniso, basis, diffusion=T)
    ip, ix, vals, N = pipeline.tensor(p2, t2, X, W, True, np.ascontiguousarray(np.moveaxis(T, (0, 1), (-2, -1))))
    check_csr(K, ip, ix, vals)


def test_tet_n1_maxwell_matches_oracle():
    import skfemntv as sk
    say("tet n1 maxwell")
    p, t = kuhn(3, 0.1, seed=5)
    mesh = sk.MeshTet(p, t)
    basis = sk.AffineTetN1Basis(mesh)
    E = t.shape[1]
    rng = np.random.default_rng(1)
    kappa = rng.uniform(0.1, 10., E)
    asm_ = sk.TetN1Assembler(basis)
    K = asm_.assemble_maxwell(mass_coefficient=kappa[:, None], curl_coefficient=1.)
    Q = len(basis.W)
    ip, ix, data, edges, ee, signs, _ = pipeline.n1_maxwell(p, t, np.repeat(kappa[:, None], Q, 1), np.ones((E, Q)))
    np.testing.assert_array_equal(basis.dof_map.topology.edges, edges)
    np.testing.assert_array_equal(basis.element_dofs, ee)
    np.testing.assert_array_equal(basis.basis_signs, signs)
    err = check_csr(K, ip, ix, data)
    say(f"n1 rel err {err:.3e}")
    # same CSR object is returned and overwritten
    K2 = asm_.assemble_mass(2.3)
    assert K2 is K
    ip, ix, data, *_ = pipeline.n1_maxwell(p, t, np.full((E, Q), 2.3), np.ones((E, Q)), include_curl=False)
    check_csr(K2, ip, ix, data)
    # a 2 KiB span budget puts 8 lanes on each edge row (incidence split):
    # the same bits as the default plan
    from skfemntv._assembly import Plan
    ref = asm_.assemble_maxwell(mass_coefficient=kappa[:, None], curl_coefficient=1.).data.copy()
    asm_.plan = Plan(basis._edges_dev, E, 6, basis.N, 1, lds_budget=2048)
    K3 = asm_.assemble_maxwell(mass_coefficient=kappa[:, None], curl_coefficient=1.)
    np.testing.assert_array_equal(K3.data, ref)


def test_assembly_is_bitwise_reproducible():
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, upload
    say("determinism")
    p, t = kuhn(4, 0.2, seed=8)
    p2, t2 = host.p2_mesh(p, t)
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=host.tet_rule_4())
    A = upload(spd_tensors(t.shape[1], 2))
    form = NativeCrossBilinearForm(basis, basis)
    a = form.assemble_tensor_device(A).clone()
    b = form.assemble_tensor_device(A).clone()
    assert bool((a == b).all())


def test_concurrent_row_split_is_reproducible():
    """Row split 3 lets the S lanes of a row add in one instruction; the LDS
    unit serialises lanes that hit one address.  Ten repeats must give the
    same bits (a wave's lanes are applied in a fixed order)."""
    import skfemntv as sk
    from skfemntv._assembly import NativeCrossBilinearForm, upload
    say("concurrent row split reproducibility")
    p, t = kuhn(10, 0.2, seed=9)
    p2, t2 = host.p2_mesh(p, t)
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=1), quadrature=host.tet_rule_4())
    form = NativeCrossBilinearForm(basis, basis)
    assert form.plan.waves > 0
    form.plan.set_tuning(row_split=3, row_form=1)
    A = upload(spd_tensors(t.shape[1], 4))
    first = form.assemble_tensor_device(A, symmetric=True).clone()
    for _ in range(9):
        assert bool((form.assemble_tensor_device(A, symmetric=True) == first).all())


def test_geometry_diagnostics_and_rejection():
    import skfemntv as sk
    say("geometry diagnostics")
    mesh = sk.MeshTet(np.array([[0., 0., 1., 0.], [0., 1., 0., 0.], [0., 0., 0., 1.]]), np.array([[0], [1], [2], [3]]))
    basis = sk.Basis(mesh, sk.ElementVector(sk.ElementTetP1(), dim=1))
    assert basis.geometry_diagnostics.maximum_determinant < 0.
    assert basis.geometry_diagnostics.negative_orientation_elements == 1
    assert basis.geometry_diagnostics.minimum_scaled_determinant == pytest.approx(1.)
    bad = sk.MeshTet(np.array([[0., 1., 0., 0.], [0., 0., 1., 1.], [0., 0., 0., 1e-30]]), np.array([[0], [1], [2], [3]]))
    with pytest.raises(ValueError, match="reason=near_singular_or_non_finite"):
        sk.Basis(bad, sk.ElementVector(sk.ElementTetP1(), dim=1))
    m2 = sk.MeshTet2.from_mesh(sk.MeshTet())
    pts = m2.p.copy()
    pts[:, 4] = -2.
    with pytest.raises(ValueError, match=r"cell=0, quadrature_point=2, determinant=-.*reason=orientation_change"):
        sk.Basis(sk.MeshTet2(pts, m2.t), sk.ElementVector(sk.ElementTetP2(), dim=1), intorder=4)


def test_curved_p2_uses_isoparametric_geometry():
    import skfemntv as sk
    from skfemntv.helpers import ddot, grad
    say("curved p2 isoparametric")
    p, t = kuhn(2, 0.0)
    p2, t2 = host.p2_mesh(p, t)
    rng = np.random.default_rng(3)
    V = p.shape[1]
    p2 = p2.copy()
    p2[:, V:] += rng.uniform(-0.02, 0.02, size=(3, p2.shape[1] - V))
    basis = sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=1))
    assert not basis._affine
    K = sk.asm(sk.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v))), basis)
    X, W = host.tet_rule_keast11()
    E = t.shape[1]
    ip, ix, vals, N = pipeline.bilinear(p2, t2, X, W, True, gradient=np.ones((E, 11)))
    check_csr(K, ip, ix, vals)


def test_tet_n1_per_point_coefficients_match_oracle():
    import skfemntv as sk
    say("tet n1 per-point coefficients")
    p, t = kuhn(2, 0.1, seed=6)
    basis = sk.AffineTetN1Basis(sk.MeshTet(p, t))
    E, Q = t.shape[1], len(basis.W)
    rng = np.random.default_rng(4)
    kq = rng.uniform(0.5, 2., (E, Q))
    mq = rng.uniform(0.5, 2., (E, Q))
    K = sk.TetN1Assembler(basis).assemble_maxwell(mass_coefficient=kq, curl_coefficient=mq)
    ip, ix, data, *_ = pipeline.n1_maxwell(p, t, kq, mq)
    check_csr(K, ip, ix, data)


def test_p1_single_element_and_larger_plan_layouts():
    """Row layouts with column splitting (long vertex rows) and tiny meshes."""
    import skfemntv as sk
    from skfemntv.helpers import ddot, grad
    for n in (1, 6):
        say(f"p1 laplace layouts n={n}")
        p, t = kuhn(n, 0.15 if n > 1 else 0.0, seed=n)
        for dim in (1, 3):
            basis = sk.Basis(sk.MeshTet(p, t), sk.ElementVector(sk.ElementTetP1(), dim=dim))
            K = sk.asm(sk.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v))), basis)
            X, W = host.tet_rule_4()
            E = t.shape[1]
            ip, ix, vals, N = pipeline.bilinear(p, t, X, W, False, components=dim, gradient=np.ones((E, 4)))
            check_csr(K, ip, ix, vals)
