"""Nodal tetrahedral elements (python/skfemntv/basis.py:926-937,1058-1061,1404-1410)."""

from __future__ import annotations

import numpy as np


class _ComposableElement:
    def __mul__(self, other):
        """``e1 * e2`` builds a composite element (basis.py:926-930)."""
        if isinstance(other, ElementComposite):
            return ElementComposite(self, *other.elems)
        return ElementComposite(self, other)


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


class ElementComposite(_ComposableElement):
    """Ordered collection of nodal H1 subfields (basis.py:1064-1076)."""

    def __init__(self, *elements):
        if len(elements) < 2:
            raise ValueError("ElementComposite requires at least two fields")
        flattened = []
        for element in elements:
            if isinstance(element, ElementComposite):
                flattened.extend(element.elems)
            else:
                flattened.append(element)
        self.elems = tuple(flattened)


__all__ = ["ElementTetP1", "ElementTetP2", "ElementVector", "ElementComposite"]
