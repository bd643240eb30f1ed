"""GPU parity of composite-basis (mixed) assembly: Taylor-Hood Stokes.

The composite numbering must equal the oracle's restatement of the
reference's node loop bit for bit; each block equals the oracle's cross
kernel on the composite DOFs (the reference's NativeCompositeBilinearForm
construction) with a bit-exact pattern and FP64 values within 1e-12
relative Frobenius.  Each case prints its name before touching the GPU.
"""

import sys

import numpy as np
import pytest
from scipy.sparse import csr_matrix

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


def taylor_hood(p2, t2, X, W):
    import skfemntv as sk
    element = sk.ElementVector(sk.ElementTetP2(), dim=3) * sk.ElementTetP1()
    return sk.Basis(sk.MeshTet2(p2, t2), element, quadrature=(X, W))


FIELDS = [(True, 3), (False, 1)]


def test_composite_numbering_matches_oracle():
    import skfemntv as sk
    say("composite numbering taylor-hood")
    p2, t2 = p2_kuhn(2, 0.2, seed=1)
    X, W = host.tet_rule_4()
    basis = taylor_hood(p2, t2, X, W)
    N, dofs = host.composite_dofs(p2.shape[1], [t2[:10], t2[:4]], [3, 1])
    assert basis.N == N
    E = t2.shape[1]
    for k, sub in enumerate(basis.subbases):
        n = dofs[k].shape[1]
        np.testing.assert_array_equal(sub.element_dofs.T.reshape(E, n, -1), dofs[k])
        assert sub.N == N
    # the default rules differ (Keast-11 for P2, 4 points for P1)
    with pytest.raises(ValueError, match="common quadrature rule"):
        sk.Basis(sk.MeshTet2(p2, t2), sk.ElementVector(sk.ElementTetP2(), dim=3) * sk.ElementTetP1())


@pytest.mark.parametrize("n", [2, 4])
def test_taylor_hood_stokes_blocks_match_oracle(n):
    import skfemntv as sk
    from skfemntv.helpers import ddot, div, grad
    p2, t2 = p2_kuhn(n, 0.2, seed=n)
    X, W = host.tet_rule_4()
    E, Q = t2.shape[1], 4
    basis = taylor_hood(p2, t2, X, W)
    native = sk.NativeCompositeBilinearForm(basis)
    nu = np.random.default_rng(n).uniform(0.5, 2., (E, 1))
    blocks = {}
    for (r, c, kind, coef) in ((0, 0, "gradient", nu), (0, 1, "row_divergence", 1.), (1, 0, "column_divergence", 1.),
                               (1, 1, "value", nu)):
        say(f"composite block ({r},{c}) {kind} n={n}")
        K = native.assemble(r, c, kind=kind, coefficient=coef)
        ip, ix, vals, N = pipeline.composite_block(p2, t2, X, W, FIELDS, r, c, kind,
                                                   coefficient=np.broadcast_to(coef, (E, Q)))
        assert K.shape == (N, N)
        np.testing.assert_array_equal(np.asarray(K.indptr, dtype=np.int64), ip)
        np.testing.assert_array_equal(np.asarray(K.indices, dtype=np.int64), ix)
        err = pipeline.relative_frobenius(K.data, vals)
        assert err <= FP64_TOL, err
        blocks[(r, c)] = csr_matrix((vals, ix, ip), shape=(N, N))

    say(f"composite stokes asm n={n}")

    @sk.BilinearForm
    def stokes(u, p, v, q, w):
        return w.nu * ddot(grad(u), grad(v)) - div(v) * p - div(u) * q

    K = sk.asm(stokes, basis, nu=nu)
    expect = blocks[(0, 0)] + (-1.) * blocks[(0, 1)] + (-1.) * blocks[(1, 0)]
    scale = np.abs(expect.data).max()
    assert abs(K - expect).max() <= 1e-12 * scale
    # the saddle-point matrix is symmetric: B^T of the row block is the column block
    assert abs(K - K.T).max() <= 1e-13 * scale
    # div(u) q annihilates nothing of the pressure, but a rigid translation
    # of the velocity has zero divergence: the (p, u) block maps it to zero
    ux = np.zeros(basis.N)
    ux[basis.subbases[0].nodal_dofs[0]] = 1.
    assert np.abs(blocks[(1, 0)] @ ux).max() <= 1e-13 * scale


def test_composite_linear_form_matches_oracle():
    import skfemntv as sk
    from skfemntv.helpers import dot
    say("composite linear form")
    p2, t2 = p2_kuhn(3, 0.2, seed=4)
    X, W = host.tet_rule_4()
    basis = taylor_hood(p2, t2, X, W)
    f = np.array([1., -2., 0.5])

    @sk.LinearForm
    def rhs(v, q, w):
        return dot(w.f, v) + q * w.g

    b = sk.asm(rhs, basis, f=f, g=np.array([3.]))
    N, cdofs = host.composite_dofs(p2.shape[1], [t2[:10], t2[:4]], [3, 1])
    expect = np.zeros(N)
    for k, ((quad, C), val) in enumerate(zip(FIELDS, (f, np.array([3.])))):
        local = pipeline.load_vector(p2, t2, X, W, quad, val, components=C)
        _, ldofs = pipeline.dofs_of(t2, p2.shape[1], 10 if quad else 4, C)
        m = np.zeros(local.shape[0], dtype=np.int64)
        m[ldofs.ravel()] = cdofs[k].ravel()
        expect[m] += local
    assert pipeline.relative_frobenius(b, expect) <= FP64_TOL
    # sum of the pressure load = 3 |Omega|
    assert abs(b[basis.subbases[1].nodal_dofs[0]].sum() - 3.) <= 1e-12
