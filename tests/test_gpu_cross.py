"""GPU parity of the cross-basis (rectangular) assembly path.

``asm(form, trial_basis, test_basis)`` and ``NativeCrossBilinearForm`` run
the HIP cross kernel; the CPU oracle's cross kernel (itself pinned
bit-exact against the reference's compiled CrossBilinearAssembler) gives
the expected CSR.  Pattern bit-exact, FP64 values within 1e-12 relative
Frobenius.  Each case prints its name before touching the GPU.
"""

import sys

import numpy as np
import pytest

from oracle import host, pipeline

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def say(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def p2_kuhn(n, jitter, seed):
    ax = np.linspace(0., 1., n + 1)
    p, t = host.kuhn_mesh(ax, ax, ax)
    if jitter:
        p = host.jitter_interior(p, 1. / n, jitter, seed)
    return host.p2_mesh(p, t)


def check_csr(K, ip, ix, vals, shape, tol=FP64_TOL):
    K = K.tocsr()
    assert K.shape == shape
    np.testing.assert_array_equal(np.asarray(K.indptr, dtype=np.int64), ip)
    np.testing.assert_array_equal(np.asarray(K.indices, dtype=np.int64), ix)
    err = pipeline.relative_frobenius(K.data, vals)
    assert err <= tol, err
    return err


def bases(p2, t2, X, W, mesh=None):
    import skfemntv as sk
    mesh = sk.MeshTet2(p2, t2) if mesh is None else mesh
    make = lambda elem, dim: sk.Basis(mesh, sk.ElementVector(elem, dim=dim), quadrature=(X, W))  # noqa: E731
    return {"p1": make(sk.ElementTetP1(), 1), "p2": make(sk.ElementTetP2(), 1),
            "vp1": make(sk.ElementTetP1(), 3), "vp2": make(sk.ElementTetP2(), 3)}


def div_tensor(c, rows_vector):
    E, Q = c.shape
    T = np.zeros((E, Q, 3, 3, 1) if rows_vector else (E, Q, 1, 3, 3))
    for k in range(3):
        if rows_vector:
            T[:, :, k, k, 0] = c
        else:
            T[:, :, 0, k, k] = c
    return T


@pytest.mark.parametrize("n,jitter", [(2, 0.2), (5, 0.15)])
def test_cross_forms_through_asm_match_oracle(n, jitter):
    import skfemntv as sk
    from skfemntv.helpers import ddot, div, dot, grad
    p2, t2 = p2_kuhn(n, jitter, seed=n)
    X, W = host.tet_rule_4()
    E, Q = t2.shape[1], 4
    B = bases(p2, t2, X, W)
    rng = np.random.default_rng(n)
    c = rng.uniform(0.5, 2., (E, 1))
    cq = np.broadcast_to(c, (E, Q))

    say(f"cross asm column divergence n={n}")
    coldiv = sk.BilinearForm(lambda u, v, w: w.c * (div(u) * v))
    K = sk.asm(coldiv, B["vp2"], B["p1"], c=c)
    ip, ix, vals, Nr, Nc = pipeline.cross(p2, t2, X, W, (False, 1), (True, 3), "value", "gradient",
                                          coefficient=div_tensor(cq, False), tensor=True)
    check_csr(K, ip, ix, vals, (B["p1"].N, B["vp2"].N))

    say(f"cross asm column divergence, per-point coefficient n={n}")
    cp = rng.uniform(0.5, 2., (E, Q))
    K = sk.asm(coldiv, B["vp2"], B["p1"], c=cp)
    _, _, vals, _, _ = pipeline.cross(p2, t2, X, W, (False, 1), (True, 3), "value", "gradient",
                                      coefficient=div_tensor(cp, False), tensor=True)
    check_csr(K, ip, ix, vals, (B["p1"].N, B["vp2"].N))

    say(f"cross asm row divergence n={n}")
    rowdiv = sk.BilinearForm(lambda u, v, w: div(v) * u)
    K = sk.asm(rowdiv, B["p1"], B["vp2"])
    ip, ix, vals, Nr, Nc = pipeline.cross(p2, t2, X, W, (True, 3), (False, 1), "gradient", "value",
                                          coefficient=div_tensor(np.ones((E, Q)), True), tensor=True)
    check_csr(K, ip, ix, vals, (B["vp2"].N, B["p1"].N))

    say(f"cross asm value + gradient n={n}")
    form = sk.BilinearForm(lambda u, v, w: w.c * dot(u, v) + 2. * ddot(grad(u), grad(v)))
    K = sk.asm(form, B["p2"], B["p1"], c=c)
    ip, ix, m_vals, Nr, Nc = pipeline.cross(p2, t2, X, W, (False, 1), (True, 1), "value", "value",
                                            coefficient=np.ascontiguousarray(cq))
    _, _, g_vals, _, _ = pipeline.cross(p2, t2, X, W, (False, 1), (True, 1), "gradient", "gradient",
                                        coefficient=np.full((E, Q), 2.))
    check_csr(K, ip, ix, m_vals + g_vals, (B["p1"].N, B["p2"].N))

    say(f"cross asm vector gradient n={n}")
    K = sk.asm(sk.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v))), B["vp1"], B["vp2"])
    ip, ix, vals, Nr, Nc = pipeline.cross(p2, t2, X, W, (True, 3), (False, 3), "gradient", "gradient")
    check_csr(K, ip, ix, vals, (B["vp2"].N, B["vp1"].N))


@pytest.mark.parametrize("case", ["tensor_vp1_p2", "tensor_p2_vp2", "value_gradient_kinds"])
def test_cross_tensor_paths_match_oracle(case):
    import skfemntv as sk
    p2, t2 = p2_kuhn(3, 0.2, seed=5)
    X, W = host.tet_rule_4()
    E, Q = t2.shape[1], 4
    B = bases(p2, t2, X, W)
    rng = np.random.default_rng(17)
    say(f"cross tensor {case}")
    if case == "tensor_vp1_p2":
        T = rng.standard_normal((E, Q, 3, 3, 1))
        native = sk.NativeCrossBilinearForm(B["vp1"], B["p2"])
        K = native.assemble_tensor("gradient", "value", T)
        ip, ix, vals, _, _ = pipeline.cross(p2, t2, X, W, (False, 3), (True, 1), "gradient", "value",
                                            coefficient=T, tensor=True)
        check_csr(K, ip, ix, vals, (B["vp1"].N, B["p2"].N))
    elif case == "tensor_p2_vp2":
        # element-constant tensor (uploaded once per element) in the full
        # gradient x gradient layout (E,Q,1,3,3,3)
        Te = rng.standard_normal((E, 1, 1, 3, 3, 3))
        native = sk.NativeCrossBilinearForm(B["p2"], B["vp2"])
        K = native.assemble_tensor("gradient", "gradient", Te)
        ip, ix, vals, _, _ = pipeline.cross(p2, t2, X, W, (True, 1), (True, 3), "gradient", "gradient",
                                            coefficient=np.ascontiguousarray(np.broadcast_to(Te, (E, Q, 1, 3, 3, 3))),
                                            tensor=True)
        check_csr(K, ip, ix, vals, (B["p2"].N, B["vp2"].N))
        # a constant tensor is uploaded as one block
        Tc = rng.standard_normal((1, 3, 3, 3))
        K = native.assemble_tensor("gradient", "gradient", Tc)
        _, _, vals, _, _ = pipeline.cross(p2, t2, X, W, (True, 1), (True, 3), "gradient", "gradient",
                                          coefficient=np.ascontiguousarray(np.broadcast_to(Tc, (E, Q, 1, 3, 3, 3))),
                                          tensor=True)
        check_csr(K, ip, ix, vals, (B["p2"].N, B["vp2"].N))
    else:
        # assemble_kinds value x gradient with a scalar: T[c,c,j] = s
        native = sk.NativeCrossBilinearForm(B["p1"], B["p2"])
        s = rng.uniform(0.5, 2., (E, Q))
        K = native.assemble_kinds("value", "gradient", s)
        T = np.zeros((E, Q, 1, 1, 3))
        T[:, :, 0, 0, :] = s[..., None]
        ip, ix, vals, _, _ = pipeline.cross(p2, t2, X, W, (False, 1), (True, 1), "value", "gradient",
                                            coefficient=T, tensor=True)
        check_csr(K, ip, ix, vals, (B["p1"].N, B["p2"].N))
        K = native.assemble_kinds("gradient", "gradient", 3.)
        _, _, vals, _, _ = pipeline.cross(p2, t2, X, W, (False, 1), (True, 1), "gradient", "gradient",
                                          coefficient=np.full((E, Q), 3.))
        check_csr(K, ip, ix, vals, (B["p1"].N, B["p2"].N))


def test_cross_isoparametric_geometry_matches_oracle():
    import skfemntv as sk
    from skfemntv.helpers import div
    say("cross curved p2 isoparametric")
    p2, t2 = p2_kuhn(2, 0.0, seed=0)
    V = (3 ** 3)
    p2 = p2.copy()
    rng = np.random.default_rng(3)
    p2[:, V:] += rng.uniform(-0.02, 0.02, size=(3, p2.shape[1] - V))
    X, W = host.tet_rule_keast11()
    B = bases(p2, t2, X, W)
    assert not B["p1"]._affine
    K = sk.asm(sk.BilinearForm(lambda u, v, w: div(u) * v), B["vp2"], B["p1"])
    E, Q = t2.shape[1], len(W)
    ip, ix, vals, _, _ = pipeline.cross(p2, t2, X, W, (False, 1), (True, 3), "value", "gradient",
                                        coefficient=div_tensor(np.ones((E, Q)), False), tensor=True)
    check_csr(K, ip, ix, vals, (B["p1"].N, B["vp2"].N))


def test_taylor_hood_divergence_closed_form_at_scale():
    """B u for the interpolants of (x,0,0), (0,0,z) and (y,0,0) on a 24^3
    jittered mesh: the P1 load vector, the load vector, zero; and the
    row-divergence block equals B transposed bit for bit.  Calls repeat
    bit-identically."""
    import skfemntv as sk
    from skfemntv.helpers import div
    say("taylor-hood closed form 24^3")
    p2, t2 = p2_kuhn(24, 0.2, seed=9)
    X, W = host.tet_rule_4()
    B = bases(p2, t2, X, W)
    Bm = sk.asm(sk.BilinearForm(lambda u, v, w: div(u) * v), B["vp2"], B["p1"])
    again = sk.asm(sk.BilinearForm(lambda u, v, w: div(u) * v), B["vp2"], B["p1"])
    np.testing.assert_array_equal(Bm.data, again.data)
    used = np.unique(t2[:10])
    x = p2[:, used]
    z = np.zeros_like(x[0])

    def vec(a, b, c):
        return np.ascontiguousarray(np.vstack([a, b, c]).T).ravel()

    (rs, rg, dx, *_), _ = pipeline.basis_geometry(p2, t2, X, W, False)
    Nr, rd = pipeline.dofs_of(t2, p2.shape[1], 4, 1)
    load = np.zeros(Nr)
    np.add.at(load, rd[:, :, 0].ravel(), np.repeat(dx.sum(axis=1) / 4., 4))
    scale = np.abs(load).max()
    np.testing.assert_allclose(Bm @ vec(x[0], z, z), load, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(Bm @ vec(z, z, x[2]), load, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(Bm @ vec(x[1], z, z), 0., rtol=0, atol=1e-12 * scale)
    Bt = sk.asm(sk.BilinearForm(lambda u, v, w: div(v) * u), B["p1"], B["vp2"])
    assert Bt.shape == (Bm.shape[1], Bm.shape[0])
    diff = (Bt - Bm.T.tocsr())
    assert diff.nnz == 0 or np.abs(diff.data).max() <= 1e-14 * np.abs(Bm.data).max()


def fan_mesh(m=120):
    """One hub vertex at the origin with m ring points and two apexes: 2m
    tetrahedra share the hub, whose P2 row then spans > 256 column nodes."""
    ang = 2 * np.pi * np.arange(m) / m
    ring = np.vstack([np.cos(ang), np.sin(ang), 0.05 * np.sin(3 * ang)])
    p = np.hstack([np.zeros((3, 1)), ring, [[0.], [0.], [1.]], [[0.], [0.], [-1.]]])
    top, bot = m + 1, m + 2
    i = np.arange(m)
    j = (i + 1) % m
    t = np.hstack([np.vstack([np.zeros(m, int), 1 + i, 1 + j, np.full(m, top)]),
                   np.vstack([np.zeros(m, int), 1 + j, 1 + i, np.full(m, bot)])])
    return p, t


def test_cross_wide_rows_use_two_byte_offsets():
    """Rows longer than 256 column nodes take 2-byte record offsets (and
    32-byte records) in both cross kernels; pattern and values against the
    oracle."""
    import skfemntv as sk
    from skfemntv.helpers import div
    say("cross wide rows (fan mesh)")
    p, t = fan_mesh(120)
    p2, t2 = host.p2_mesh(p, t)
    X, W = host.tet_rule_4()
    B = bases(p2, t2, X, W)
    native = sk.NativeCrossBilinearForm(B["p1"], B["vp2"])
    assert native.cross_plan().max_row_len > 256
    E, Q = t2.shape[1], 4
    for coef, name in ((np.ones((E, 1)), "element-constant"), (np.random.default_rng(2).uniform(.5, 2., (E, Q)),
                                                              "per-point")):
        say(f"cross wide rows {name}")
        K = sk.asm(sk.BilinearForm(lambda u, v, w: w.c * (div(u) * v)), B["vp2"], B["p1"], c=coef)
        ip, ix, vals, _, _ = pipeline.cross(p2, t2, X, W, (False, 1), (True, 3), "value", "gradient",
                                            coefficient=div_tensor(np.broadcast_to(coef, (E, Q)), False),
                                            tensor=True)
        check_csr(K, ip, ix, vals, (B["p1"].N, B["vp2"].N))
