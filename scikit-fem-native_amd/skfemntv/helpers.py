"""Form helpers (python/skfemntv/helpers.py:1-22)."""

import numpy as np

from ._symbols import apply_symgrad
from .forms import ddot, div, dot, grad, mul


def trace(value):
    return np.einsum("ii...->...", value)


def transpose(value):
    return np.einsum("ij...->ji...", value)


def sym_grad(value):
    return apply_symgrad(value)


def jump(w, *args):
    """Two-sided jump on interior facets (interface.py:20-33): with the pair
    of sides w.idx = (trial side, test side) the k-th argument is multiplied
    by (-1) ** w.idx[k]."""
    if not hasattr(w, "idx"):
        raise TypeError("jump expects the parameters of an interior-facet form (w.idx)")
    values = tuple(((-1.) ** w.idx[k]) * value for k, value in enumerate(args))
    return values[0] if len(values) == 1 else values


__all__ = ["dot", "ddot", "div", "grad", "jump", "mul", "sym_grad", "trace", "transpose"]
