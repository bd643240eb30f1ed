"""Form helpers (python/skfemntv/helpers.py:1-22)."""

import numpy as np

from ._symbols import SymmetricGradient, TestGradient, TrialGradient
from .forms import ddot, div, dot, grad, mul


def trace(value):
    return np.einsum("ii...->...", value)


def transpose(value):
    return np.einsum("ij...->ji...", value)


def sym_grad(value):
    gradient = grad(value)
    if isinstance(gradient, (TrialGradient, TestGradient)):
        role = "trial" if isinstance(gradient, TrialGradient) else "test"
        return SymmetricGradient(role, gradient.factor)
    return 0.5 * (gradient + transpose(gradient))


__all__ = ["dot", "ddot", "div", "grad", "mul", "sym_grad", "trace", "transpose"]
