"""Typed forms and ``asm`` (python/skfemntv/forms.py:71-112,161-404,667-1091).

``asm(form, basis, **coefficients)`` traces the user function with symbolic
trial/test nodes into the reference's term kinds, evaluates coefficients on
the host exactly as the reference does, and assembles on the GPU.  Forms
the MI355X build does not assemble raise ``UnsupportedNativeForm``; nothing
falls back to a CPU path.
"""

from __future__ import annotations

from typing import Callable
from weakref import WeakKeyDictionary

import numpy as np

from ._assembly import (NativeBilinearForm, NativeCompositeBilinearForm, NativeCompositeLinearForm,
                        NativeCrossBilinearForm, NativeLinearForm)
from ._symbols import (TEST, TRIAL, Field, PointValue, UnsupportedNativeForm, apply_div, apply_grad, apply_mul,
                       coefficient_array, collect, contract, namespace, resolve, trace)


class _LinearForm:
    def __init__(self, function: Callable):
        self.function = function
        self._native_cache = WeakKeyDictionary()

    def assemble(self, *bases, **kwargs):
        return asm(self, *bases, **kwargs)


class _BilinearForm:
    def __init__(self, function: Callable):
        self.function = function
        self._native_cache = WeakKeyDictionary()
        self._cross_cache = WeakKeyDictionary()

    def assemble(self, *bases, **kwargs):
        return asm(self, *bases, **kwargs)


class _Functional:
    def __init__(self, function: Callable):
        self.function = function

    def assemble(self, *bases, **kwargs):
        return asm(self, *bases, **kwargs)


def LinearForm(function=None, **kwargs):
    if kwargs:
        return lambda fn: _LinearForm(fn)
    return _LinearForm(function) if function is not None else _LinearForm


def Functional(function=None, **kwargs):
    """A scalar integrand of the geometry and coefficients (forms.py:77-106):
    ``asm(functional, basis)`` returns its integral."""
    if kwargs:
        return lambda fn: _Functional(fn)
    return _Functional(function) if function is not None else _Functional


def _native_functional_assemble(form, basis, kwargs):
    """forms.py (``_native_functional_assemble``): the integrand is evaluated
    at the quadrature points, then sum(values * dx) is reduced on the GPU
    (integrate_functional, functional_assembler.cpp:12-33)."""
    from ._assembly import integrate_functional
    try:
        expression = form.function(namespace(_geometry(basis), kwargs))
        values = np.asarray(resolve(expression, kwargs), dtype=np.float64)
    except UnsupportedNativeForm:
        raise
    except Exception as error:
        raise UnsupportedNativeForm(f"Functional contains an unsupported operation: {error}") from error
    try:
        values = np.broadcast_to(values, basis.dx.shape)
    except ValueError as error:
        raise UnsupportedNativeForm("Functional must evaluate to one scalar per quadrature point") from error
    return integrate_functional(values, basis.dx)


def BilinearForm(function=None, **kwargs):
    if kwargs:
        return lambda fn: _BilinearForm(fn)
    return _BilinearForm(function) if function is not None else _BilinearForm


def _geometry(basis):
    # Physical quadrature coordinates, component first; computed on the GPU
    # only when the form reads w.x.  Facet bases add their outward normals
    # as w.n (forms.py:418-421).
    geometry = {"x": PointValue(lambda: np.moveaxis(basis.global_coordinates, -1, 0))}
    if getattr(basis, "normals", None) is not None:
        geometry["n"] = PointValue(np.moveaxis(basis.normals, -1, 0))
    return geometry


_LINEAR_MESSAGE = ("native LinearForm must reduce to dot(coefficient, v), ddot(coefficient, grad(v)), "
                   "or a sum of those terms")
_BILINEAR_MESSAGE = "native BilinearForm must reduce to dot(u, v), ddot(grad(u), grad(v)), or a sum of those terms"


def _compile_linear(form, basis, kwargs):
    expression = trace(form.function, (Field(TEST),), _geometry(basis), kwargs, context="LinearForm")
    return collect(expression, linear=True, message=_LINEAR_MESSAGE)


def _native_linear_assemble(form, basis, kwargs, num_threads=0):
    """forms.py:161-203."""
    terms = _compile_linear(form, basis, kwargs)
    native = form._native_cache.get(basis)
    if native is None:
        from .facet import FacetBasis, NativeFacetLinearForm
        native = NativeFacetLinearForm(basis) if isinstance(basis, FacetBasis) else NativeLinearForm(basis)
        form._native_cache[basis] = native
    value = None
    gradient = None
    for term in terms:
        coefficient = coefficient_array(term, kwargs)
        if term.kind == "value":
            if native._shape[-1] == 1 and np.shape(coefficient) == native._shape[:2]:
                coefficient = np.asarray(coefficient)[..., None]
            elif np.ndim(coefficient) and np.shape(coefficient)[0] == native._shape[-1]:
                coefficient = np.moveaxis(coefficient, 0, -1)
            value = coefficient if value is None else value + coefficient
        elif term.kind == "gradient":
            if np.ndim(coefficient) >= 2 and np.shape(coefficient)[:2] == native._gradient_shape[-2:]:
                coefficient = np.moveaxis(coefficient, (0, 1), (-2, -1))
            gradient = coefficient if gradient is None else gradient + coefficient
        else:
            raise UnsupportedNativeForm
    result, _ = native.assemble(value=value, gradient=gradient, num_threads=num_threads)
    if result.shape[0] == basis.N:
        return result
    padded = np.zeros(basis.N, dtype=result.dtype)
    padded[:result.shape[0]] = result
    return padded


def anisotropic_tensor_coefficient(raw, coefficient_shape, dimension):
    """Normalise a diffusion tensor to (E,Q,1,dim,1,dim) (forms.py:377-404):
    accepts (dim,dim), the recommended (dim,dim,E,Q) or the native (E,Q,dim,dim)."""
    tensor = np.asarray(raw, dtype=np.float64)
    coefficient_shape = tuple(coefficient_shape)
    if tensor.shape == (dimension, dimension):
        tensor = np.broadcast_to(tensor, coefficient_shape + (dimension, dimension))
    elif tensor.shape == (dimension, dimension) + coefficient_shape:
        tensor = np.moveaxis(tensor, (0, 1), (-2, -1))
    elif tensor.shape == coefficient_shape + (dimension, dimension):
        pass
    else:
        raise ValueError(
            "anisotropic tensor must use the recommended scikit-fem shape "
            f"{(dimension, dimension) + coefficient_shape}, constant shape ({dimension}, {dimension}), "
            f"or the low-level native compatibility shape {coefficient_shape + (dimension, dimension)}; "
            f"got {tensor.shape}")
    return tensor[..., None, :, None, :]


def _native_bilinear_assemble(form, basis, kwargs, num_threads=0, *, memory_limit_bytes=None,
                              memory_safety_factor=1.25, precision="fp64"):
    """forms.py:276-374."""
    expression = trace(form.function, (Field(TRIAL), Field(TEST)), _geometry(basis), kwargs, context="BilinearForm")
    terms = collect(expression, linear=False, message=_BILINEAR_MESSAGE)
    kinds = {"value": None, "gradient": None, "symmetric_gradient": None, "divergence": None}
    tensor_terms = []
    for term in terms:
        coefficient = coefficient_array(term, kwargs, squeeze=True)
        if term.kind in kinds:
            prev = kinds[term.kind]
            kinds[term.kind] = coefficient if prev is None else prev + coefficient
        elif term.kind == "gradient_tensor":
            tensor_terms.append((coefficient, 1.))
        else:
            raise UnsupportedNativeForm(f"unsupported bilinear term kind {term.kind!r}")
    result = None
    facet = hasattr(basis, "parent_elements")
    if facet and tensor_terms:
        raise UnsupportedNativeForm("facet bilinear forms take the value kind in the MI355X build")
    if any(v is not None for v in kinds.values()):
        native = form._native_cache.get(basis)
        if native is None:
            if facet:
                from .facet import NativeFacetBilinearForm
                native = NativeFacetBilinearForm(basis)
            else:
                native = NativeBilinearForm(basis, memory_limit_bytes=memory_limit_bytes,
                                            memory_safety_factor=memory_safety_factor)
            form._native_cache[basis] = native
        result = native.assemble(num_threads=num_threads, precision=precision, **kinds)
    if tensor_terms:
        if basis.elem._dim != 1:
            raise UnsupportedNativeForm("anisotropic gradient tensor currently requires a scalar field")
        by_trial = form._cross_cache.setdefault(basis, WeakKeyDictionary())
        tensor_native = by_trial.get(basis)
        if tensor_native is None:
            tensor_native = NativeCrossBilinearForm(basis, basis, memory_limit_bytes=memory_limit_bytes,
                                                    memory_safety_factor=memory_safety_factor)
            by_trial[basis] = tensor_native
        for raw, factor in tensor_terms:
            tensor = anisotropic_tensor_coefficient(raw, basis.coefficient_shape, 3)
            if factor != 1.0:
                tensor = tensor * factor
            matrix = tensor_native.assemble_tensor("gradient", "gradient", tensor, num_threads=num_threads,
                                                   precision=precision)
            result = matrix if result is None else result + matrix
    if result is None:
        raise UnsupportedNativeForm("bilinear form contains no assembled terms")
    return result


def _native_interior_bilinear_assemble(form, trial_bases, test_bases, kwargs):
    """forms.py:471-487: ``asm(form, [trial sides], [test sides])`` sums the
    cross-basis form over every (trial side, test side) pair; the form reads
    the pair as ``w.idx`` = (trial index, test index)."""
    trial_bases, test_bases = tuple(trial_bases), tuple(test_bases)
    if not trial_bases or not test_bases:
        raise ValueError("interior facet basis lists cannot be empty")
    result = None
    for i, trial in enumerate(trial_bases):
        for j, test in enumerate(test_bases):
            matrix = _native_cross_bilinear_assemble(form, trial, test, kwargs, idx=(i, j))
            result = matrix if result is None else result + matrix
    return result


def _native_cross_bilinear_assemble(form, trial_basis, test_basis, kwargs, *, memory_limit_bytes=None,
                                    memory_safety_factor=1.25, idx=None):
    """forms.py:407-468: ``asm(form, trial_basis, test_basis)``; rows follow
    the test basis, columns the trial basis, one fresh CSR per call."""
    geometry = _geometry(test_basis)
    if idx is not None:
        geometry["idx"] = idx
    expression = trace(form.function, (Field(TRIAL), Field(TEST)), geometry, kwargs,
                       context="cross-basis BilinearForm")
    terms = collect(expression, linear=False,
                    message="cross-basis BilinearForm must reduce to value, gradient, or divergence contractions")
    by_trial = form._cross_cache.setdefault(trial_basis, WeakKeyDictionary())
    native = by_trial.get(test_basis)
    if native is None:
        native = NativeCrossBilinearForm(test_basis, trial_basis, memory_limit_bytes=memory_limit_bytes,
                                         memory_safety_factor=memory_safety_factor)
        by_trial[test_basis] = native
    elif memory_limit_bytes is not None:
        from ._assembly import enforce_memory_budget
        enforce_memory_budget(native.memory_estimate, memory_limit_bytes, memory_safety_factor)
    result = None
    for term in terms:
        if term.kind not in ("value", "gradient", "row_divergence", "column_divergence"):
            raise UnsupportedNativeForm(f"cross-basis BilinearForm does not support {term.kind!r} terms")
        coefficient = coefficient_array(term, kwargs, squeeze=True)
        matrix = native.assemble(term.kind, coefficient)
        result = matrix if result is None else result + matrix
    return result


def _native_composite_bilinear_assemble(form, basis, kwargs):
    """forms.py:489-539: the form takes the trial subfields, then the test
    subfields; every term is one (test field, trial field) block."""
    fields = len(basis.subbases)
    args = tuple(Field(TRIAL, k) for k in range(fields)) + tuple(Field(TEST, k) for k in range(fields))
    expression = trace(form.function, args, _geometry(basis), kwargs, context="composite BilinearForm")
    terms = collect(expression, linear=False,
                    message="composite BilinearForm must reduce to subfield value or gradient contractions")
    native = form._native_cache.get(basis)
    if native is None:
        native = NativeCompositeBilinearForm(basis)
        form._native_cache[basis] = native
    result = None
    for term in terms:
        coefficient = coefficient_array(term, kwargs, squeeze=True)
        block = native.assemble(term.test_sub, term.trial_sub, kind=term.kind, coefficient=coefficient)
        result = block if result is None else result + block
    return result


def _native_composite_linear_assemble(form, basis, kwargs):
    """forms.py:206-273: value / gradient terms per test subfield."""
    test = tuple(Field(TEST, k) for k in range(len(basis.subbases)))
    expression = trace(form.function, test, _geometry(basis), kwargs, context="composite LinearForm")
    terms = collect(expression, linear=True,
                    message="composite LinearForm must reduce to subfield value or gradient contractions")
    native = form._native_cache.get(basis)
    if native is None:
        native = NativeCompositeLinearForm(basis)
        form._native_cache[basis] = native
    grouped = {}
    for term in terms:
        coefficient = coefficient_array(term, kwargs)
        field_native = native.assembler(term.test_sub)
        if term.kind == "value":
            if field_native._shape[-1] == 1 and np.shape(coefficient) == field_native._shape[:2]:
                coefficient = np.asarray(coefficient)[..., None]
            elif np.ndim(coefficient) and np.shape(coefficient)[0] == field_native._shape[-1]:
                coefficient = np.moveaxis(coefficient, 0, -1)
        elif term.kind == "gradient":
            if np.ndim(coefficient) >= 2 and np.shape(coefficient)[:2] == field_native._gradient_shape[-2:]:
                coefficient = np.moveaxis(coefficient, (0, 1), (-2, -1))
        else:
            raise UnsupportedNativeForm(f"unsupported composite linear kind {term.kind!r}")
        values = grouped.setdefault(term.test_sub, {"value": None, "gradient": None})
        previous = values[term.kind]
        values[term.kind] = coefficient if previous is None else previous + coefficient
    result = np.zeros(basis.N, dtype=np.float64)
    for field, values in grouped.items():
        result += native.assemble(field, **values)
    return result


def asm(form, *bases, num_threads=None, memory_limit_bytes=None, memory_safety_factor=1.25, **kwargs):
    """Assemble strictly with the native (MI355X) backend.

    Unsupported forms raise ``UnsupportedNativeForm``; this function never
    delegates to a CPU implementation.
    """
    requested = 0
    if num_threads is not None:
        from .runtime import available_num_threads
        if isinstance(num_threads, bool) or not isinstance(num_threads, int) or num_threads < 1:
            raise ValueError("num_threads must be a positive integer")
        requested = min(num_threads, available_num_threads())
    if isinstance(form, _Functional):
        if kwargs.pop("integration", None) is not None:
            raise UnsupportedNativeForm("interface assembly is outside the MI355X build")
        if len(bases) != 1:
            raise UnsupportedNativeForm("native Functional requires one Basis or FacetBasis")
        return _native_functional_assemble(form, bases[0], kwargs)
    if isinstance(form, _LinearForm):
        if kwargs.pop("integration", None) is not None:
            raise UnsupportedNativeForm("interface assembly is outside the MI355X build")
        if len(bases) == 1:
            if hasattr(bases[0], "subbases"):
                if num_threads is not None:
                    raise UnsupportedNativeForm("per-call threads are not yet supported for composite LinearForm")
                return _native_composite_linear_assemble(form, bases[0], kwargs)
            _reject_composite_member(bases[0])
            return _native_linear_assemble(form, bases[0], kwargs, requested)
        raise UnsupportedNativeForm("native LinearForm requires one basis, or two bases with an interface "
                                    "integration")
    if isinstance(form, _BilinearForm):
        if kwargs.pop("integration", None) is not None:
            raise UnsupportedNativeForm("interface assembly is outside the MI355X build")
        if len(bases) == 2 and all(isinstance(b, (list, tuple)) for b in bases):
            if num_threads is not None:
                raise UnsupportedNativeForm("per-call threads are not yet supported for interior BilinearForm")
            return _native_interior_bilinear_assemble(form, bases[0], bases[1], kwargs)
        if len(bases) == 2:
            if any(isinstance(b, (list, tuple)) for b in bases):
                raise UnsupportedNativeForm("interior facet assembly takes two lists of facet bases")
            if num_threads is not None:
                raise UnsupportedNativeForm("per-call threads are not yet supported for cross-basis BilinearForm")
            for b in bases:
                _reject_composite_member(b)
            return _native_cross_bilinear_assemble(form, bases[0], bases[1], kwargs,
                                                   memory_limit_bytes=memory_limit_bytes,
                                                   memory_safety_factor=memory_safety_factor)
        if len(bases) != 1:
            raise UnsupportedNativeForm("native BilinearForm currently requires one shared basis")
        if hasattr(bases[0], "subbases"):
            if num_threads is not None:
                raise UnsupportedNativeForm("per-call threads are not yet supported for composite BilinearForm")
            return _native_composite_bilinear_assemble(form, bases[0], kwargs)
        _reject_composite_member(bases[0])
        return _native_bilinear_assemble(form, bases[0], kwargs, requested, memory_limit_bytes=memory_limit_bytes,
                                         memory_safety_factor=memory_safety_factor)
    raise TypeError("skfemntv.asm accepts forms created by skfemntv.LinearForm or skfemntv.BilinearForm")


def trace_terms(form, kwargs=None, *, geometry=None, fields=0):
    """Trace a form without assembling it and list its terms as tuples
    (kind, factor, coefficient, row field, column field); coefficients are
    arrays, or the parameter name of a coefficient not supplied in kwargs;
    fields > 0 traces a composite form with that many subfields (row =
    test field, column = trial field; a linear term has its field as the
    row).  Errors are those of asm."""
    kwargs = dict(kwargs or {})
    geometry = {k: (PointValue(v) if isinstance(v, np.ndarray) else v) for k, v in (geometry or {}).items()}
    bilinear = isinstance(form, _BilinearForm)
    if fields:
        test = tuple(Field(TEST, k) for k in range(fields))
        args = tuple(Field(TRIAL, k) for k in range(fields)) + test if bilinear else test
    else:
        args = (Field(TRIAL), Field(TEST)) if bilinear else (Field(TEST),)
    expression = trace(form.function, args, geometry, kwargs, context="form")
    terms = collect(expression, linear=not bilinear, message=_BILINEAR_MESSAGE if bilinear else _LINEAR_MESSAGE)

    def coef(c):
        if c is None:
            return None
        if hasattr(c, "name") and hasattr(c, "index") and not isinstance(c, np.ndarray):
            return c.name if c.index is None else (c.name, c.index)
        return np.asarray(c)

    return [(t.kind, t.scale, coef(t.coef), t.test_sub, t.trial_sub) for t in terms]


def _reject_composite_member(basis):
    # a sub-basis of a composite basis carries composite DOF numbers; its
    # blocks are assembled through the composite basis
    if getattr(basis, "_composite_parent", None) is not None:
        raise UnsupportedNativeForm("assemble composite subfields through their composite basis")


# ------------------------------------------------------------------ operators
# With a symbolic trial or test field among the arguments these build the
# form's terms; with numbers only they are the numeric contractions
# (scikit-fem's helpers) over the leading component axes.
def dot(left, right):
    if isinstance(left, Field) or isinstance(right, Field):
        return contract(left, right, "dot")
    return np.einsum("i...,i...->...", left, right)


def ddot(left, right):
    if isinstance(left, Field) or isinstance(right, Field):
        return contract(left, right, "ddot")
    return np.einsum("ij...,ij...->...", left, right)


def mul(left, right):
    if isinstance(left, Field) or isinstance(right, Field):
        return apply_mul(left, right)
    return np.einsum("ij...,j...->i...", left, right)


def grad(value):
    return apply_grad(value)


def div(value):
    return apply_div(value)


__all__ = ["BilinearForm", "LinearForm", "asm", "dot", "ddot", "mul", "grad", "div", "UnsupportedNativeForm",
           "anisotropic_tensor_coefficient"]
