"""Nodal tetrahedral elements (python/skfemntv/basis.py:926-937,1058-1061,1404-1410)."""

from __future__ import annotations

import numpy as np


class _ComposableElement:
    def __mul__(self, other):
        raise NotImplementedError("composite (mixed) elements are outside the MI355X assembly build")


class ElementTetP1(_ComposableElement):
    def __init__(self):
        self.doflocs = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.]])


class ElementTetP2(_ComposableElement):
    def __init__(self):
        self.doflocs = np.array(
            [[0., 0., 0.], [1., 0., 0.], [0., 1., 0.], [0., 0., 1.],
             [.5, 0., 0.], [.5, .5, 0.], [0., .5, 0.], [0., 0., .5],
             [.5, 0., .5], [0., .5, .5]])


class ElementVector(_ComposableElement):
    def __init__(self, element, dim: int | None = None):
        self.elem = element
        self._dim = element.doflocs.shape[1] if dim is None else dim


__all__ = ["ElementTetP1", "ElementTetP2", "ElementVector"]
