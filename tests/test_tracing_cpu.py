"""Form tracing on the host: which term kinds, factors and coefficients a
user form reduces to, and which forms are rejected with which message.

The kinds are the reference's (python/skfemntv/_form_terms.py: value,
gradient, symmetric_gradient, divergence, row_divergence,
column_divergence, gradient_tensor); the rejections keep its messages
(_h1_fields.py:62-71 "weighted divergence products", forms.py:851-961).
No device is needed: tracing only builds the term list that asm assembles.
"""

import numpy as np
import pytest

import skfemntv as sk
from skfemntv import UnsupportedNativeForm
from skfemntv.forms import trace_terms
from skfemntv.helpers import ddot, div, dot, grad, mul, sym_grad, trace, transpose

E, Q = 3, 4
rng = np.random.default_rng(0)
K = rng.uniform(1, 2, (E, Q))
MU = rng.uniform(1, 2, (E, 1))
LAM = rng.uniform(1, 2, (E, 1))
A = rng.standard_normal((3, 3, E, Q))
F = rng.standard_normal((E, Q))
G = rng.standard_normal((3, E, Q))
X = rng.standard_normal((3, E, Q))


def kinds(terms):
    return [(t[0], t[1]) for t in terms]


def test_bilinear_kinds_and_factors():
    cases = [
        (lambda u, v, w: dot(u, v), [("value", 1.)]),
        (lambda u, v, w: 3. * dot(u, v), [("value", 3.)]),
        (lambda u, v, w: dot(2. * u, v), [("value", 2.)]),
        (lambda u, v, w: -dot(u, v), [("value", -1.)]),
        (lambda u, v, w: ddot(grad(u), grad(v)), [("gradient", 1.)]),
        (lambda u, v, w: dot(grad(u), grad(v)), [("gradient", 1.)]),
        (lambda u, v, w: ddot(grad(v), grad(u)), [("gradient", 1.)]),
        (lambda u, v, w: dot(u, v) - ddot(grad(u), grad(v)), [("value", 1.), ("gradient", -1.)]),
        (lambda u, v, w: ddot(sym_grad(u), sym_grad(v)), [("symmetric_gradient", 1.)]),
        (lambda u, v, w: div(u) * div(v), [("divergence", 1.)]),
        (lambda u, v, w: div(u) * v, [("column_divergence", 1.)]),
        (lambda u, v, w: div(v) * u, [("row_divergence", 1.)]),
        (lambda u, v, w: v * div(u), [("column_divergence", 1.)]),
        (lambda u, v, w: dot(u, v) + dot(u, v) + ddot(grad(u), grad(v)),
         [("value", 1.), ("value", 1.), ("gradient", 1.)]),
    ]
    for fn, want in cases:
        assert kinds(trace_terms(sk.BilinearForm(fn))) == want


def test_bilinear_coefficients():
    t = trace_terms(sk.BilinearForm(lambda u, v, w: w.k * ddot(grad(u), grad(v))), {"k": K})
    assert kinds(t) == [("gradient", 1.)]
    np.testing.assert_array_equal(t[0][2], K)
    t = trace_terms(sk.BilinearForm(lambda u, v, w: ddot(grad(u), grad(v)) * w.k), {"k": K})
    np.testing.assert_array_equal(t[0][2], K)
    t = trace_terms(sk.BilinearForm(
        lambda u, v, w: 2. * w.mu * ddot(sym_grad(u), sym_grad(v)) + w.lam * (div(u) * div(v))),
        {"mu": MU, "lam": LAM})
    assert kinds(t) == [("symmetric_gradient", 1.), ("divergence", 1.)]
    np.testing.assert_array_equal(t[0][2], 2. * MU)
    np.testing.assert_array_equal(t[1][2], LAM)
    t = trace_terms(sk.BilinearForm(lambda u, v, w: dot(mul(w.A, grad(u)), grad(v))), {"A": A})
    assert kinds(t) == [("gradient_tensor", 1.)]
    np.testing.assert_array_equal(t[0][2], A)
    t = trace_terms(sk.BilinearForm(lambda u, v, w: dot(grad(v), mul(w.A, grad(u)))), {"A": A})
    assert kinds(t) == [("gradient_tensor", 1.)]
    t = trace_terms(sk.BilinearForm(lambda u, v, w: w.c * div(u) * v), {"c": F})
    assert kinds(t) == [("column_divergence", 1.)]
    np.testing.assert_array_equal(t[0][2], F)
    # a coefficient that is not supplied stays symbolic until asm evaluates it
    t = trace_terms(sk.BilinearForm(lambda u, v, w: w.missing * dot(u, v)))
    assert t[0][:2] == ("value", 1.) and t[0][2] == "missing"
    # geometry values trace like coefficients
    t = trace_terms(sk.BilinearForm(lambda u, v, w: w.x[0] * dot(u, v)), geometry={"x": X})
    np.testing.assert_array_equal(t[0][2], X[0])


@pytest.mark.parametrize("fn, message", [
    (lambda u, v, w: w.lam * div(u) * div(v), "weighted divergence products"),
    (lambda u, v, w: dot(mul(w.A, grad(u)), grad(u)), "anisotropic contraction requires trial and test"),
    (lambda u, v, w: dot(u, v) * w.k * w.k, "multiple bilinear coefficients"),
    (lambda u, v, w: dot(u, grad(v)), "unsupported dot"),
    (lambda u, v, w: mul(u, grad(v)), "native mul"),
    (lambda u, v, w: u * v, "unsupported operation"),
])
def test_bilinear_rejections(fn, message):
    with pytest.raises(UnsupportedNativeForm, match=message):
        trace_terms(sk.BilinearForm(fn), {"lam": LAM, "A": A, "k": K})


def test_linear_kinds_and_coefficients():
    t = trace_terms(sk.LinearForm(lambda v, w: dot(w.f, v)), {"f": F})
    assert kinds(t) == [("value", 1.)]
    np.testing.assert_array_equal(t[0][2], F)
    t = trace_terms(sk.LinearForm(lambda v, w: dot(v, w.f)), {"f": F})
    assert kinds(t) == [("value", 1.)]
    t = trace_terms(sk.LinearForm(lambda v, w: ddot(w.G, grad(v))), {"G": G})
    assert kinds(t) == [("gradient", 1.)]
    np.testing.assert_array_equal(t[0][2], G)
    t = trace_terms(sk.LinearForm(lambda v, w: dot(w.f, v) - 2. * ddot(w.G, grad(v))), {"f": F, "G": G})
    assert kinds(t) == [("value", 1.), ("gradient", -2.)]
    t = trace_terms(sk.LinearForm(lambda v, w: -dot(w.f, v)), {"f": F})
    assert kinds(t) == [("value", -1.)]
    t = trace_terms(sk.LinearForm(lambda v, w: dot(w.g, v)))
    assert t[0][:3] == ("value", 1., "g")
    t = trace_terms(sk.LinearForm(lambda v, w: dot(w.x[0] * w.x[1], v)), geometry={"x": X})
    np.testing.assert_allclose(t[0][2], X[0] * X[1])
    with pytest.raises(UnsupportedNativeForm, match="must reduce to"):
        trace_terms(sk.LinearForm(lambda v, w: w.f * v), {"f": F})


def test_composite_blocks():
    def stokes(u, p, v, q, w):
        return w.nu * ddot(grad(u), grad(v)) - div(v) * p - div(u) * q + 2. * (p * q)

    t = trace_terms(sk.BilinearForm(stokes), {"nu": K}, fields=2)
    assert [(x[0], x[1], x[3], x[4]) for x in t] == [
        ("gradient", 1., 0, 0), ("row_divergence", -1., 0, 1), ("column_divergence", -1., 1, 0),
        ("value", 2., 1, 1)]
    np.testing.assert_array_equal(t[0][2], K)
    t = trace_terms(sk.BilinearForm(lambda u, p, v, q, w: dot(u, v) + w.c * (p * q)), {"c": F}, fields=2)
    assert [(x[0], x[3], x[4]) for x in t] == [("value", 0, 0), ("value", 1, 1)]
    np.testing.assert_array_equal(t[1][2], F)
    t = trace_terms(sk.LinearForm(lambda v, q, w: dot(w.f3, v) + q * w.g), {"f3": G, "g": F}, fields=2)
    assert [(x[0], x[3]) for x in t] == [("value", 0), ("value", 1)]
    np.testing.assert_array_equal(t[0][2], G)
    np.testing.assert_array_equal(t[1][2], F)
    t = trace_terms(sk.LinearForm(lambda v, q, w: ddot(w.H, grad(v))), {"H": A}, fields=2)
    assert [(x[0], x[3]) for x in t] == [("gradient", 0)]


@pytest.mark.parametrize("fn, message", [
    (lambda u, p, v, q, w: ddot(grad(u), grad(p)), "requires trial and test subfields"),
    (lambda u, p, v, q, w: dot(grad(u), v), "composite dot requires two values or two gradients"),
    (lambda u, p, v, q, w: ddot(u, v), "composite ddot requires two gradients"),
    (lambda u, p, v, q, w: div(grad(u)) * q, "div\\(\\) expects a composite value"),
    (lambda u, p, v, q, w: div(u) * p, "divergence coupling requires trial and test subfields"),
])
def test_composite_rejections(fn, message):
    with pytest.raises(UnsupportedNativeForm, match=message):
        trace_terms(sk.BilinearForm(fn), fields=2)


def test_composite_linear_rejects_other_operators():
    # a linear form reduces to value or gradient contractions of test subfields
    with pytest.raises(UnsupportedNativeForm):
        trace_terms(sk.LinearForm(lambda v, q, w: div(v) * w.g), {"g": F}, fields=2)


def test_numeric_helpers_stay_numeric():
    M = rng.standard_normal((3, 3, E))
    np.testing.assert_allclose(trace(M), M[0, 0] + M[1, 1] + M[2, 2])
    np.testing.assert_array_equal(transpose(M), np.swapaxes(M, 0, 1))
    np.testing.assert_allclose(ddot(M, M), (M * M).sum(axis=(0, 1)))
    np.testing.assert_allclose(dot(G, G), (G * G).sum(axis=0))
