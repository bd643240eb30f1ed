"""Symbolic trial/test nodes, typed terms and coefficient values used when a
user form function is traced.

These restate the typed-form semantics of skfemntv 0.3.1
(python/skfemntv/_h1_fields.py:14-89, _form_terms.py:14-113,
_coefficients.py:15-86, _form_parameters.py:14-121,
_form_compiler.py:9-30, _errors.py:4-5) so that forms written for the
reference trace to the same term kinds and coefficient arrays here.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Mapping

import numpy as np


class UnsupportedNativeForm(Exception):
    """Raised when a form cannot be assembled by the native backend."""


# ------------------------------------------------------------------ coefficients
@dataclass(frozen=True)
class Coefficient:
    """A named form parameter that has not been supplied yet."""
    name: str

    def __getitem__(self, index):
        if isinstance(index, bool) or not isinstance(index, (int, np.integer)):
            raise ValueError("coefficient component index must be an integer")
        return CoefficientComponent(self.name, int(index))


@dataclass(frozen=True)
class CoefficientComponent:
    name: str
    index: int


def is_symbolic(value) -> bool:
    return isinstance(value, (Coefficient, CoefficientComponent))


def resolve_coefficient(value: Any, parameters: Mapping[str, Any]) -> Any:
    if isinstance(value, Coefficient):
        name, index = value.name, None
    elif isinstance(value, CoefficientComponent):
        name, index = value.name, value.index
    elif isinstance(value, str):
        name, index = value, None
    else:
        return value
    if name not in parameters:
        raise ValueError(f"missing form parameter {name!r}")
    resolved = parameters[name]
    if index is None:
        return resolved
    array = np.asarray(resolved)
    try:
        return array[index]
    except IndexError as error:
        raise ValueError(f"form parameter {name!r} has no component {index}") from error


def evaluate_coefficient(value, parameters, *, factor=1.0, squeeze=False):
    """None -> the bare factor; otherwise factor * float64 array, squeezed
    when asked and ndim > 2 (_coefficients.py:71-86)."""
    if value is None:
        return factor
    evaluated = np.asarray(resolve_coefficient(value, parameters), dtype=np.float64)
    if factor != 1.0:   # x * 1.0 == x exactly; keep broadcast views as views
        evaluated = factor * evaluated
    if squeeze and evaluated.ndim > 2:
        evaluated = np.squeeze(evaluated)
    return evaluated


# ------------------------------------------------------------------ fields
@dataclass(frozen=True)
class TestValue:
    factor: float = 1.0

    def __mul__(self, value):
        return TestValue(self.factor * value) if np.isscalar(value) else NotImplemented

    __rmul__ = __mul__


@dataclass(frozen=True)
class TestGradient:
    factor: float = 1.0


@dataclass(frozen=True)
class TrialValue:
    factor: float = 1.0

    def __mul__(self, value):
        return TrialValue(self.factor * value) if np.isscalar(value) else NotImplemented

    __rmul__ = __mul__


@dataclass(frozen=True)
class TrialGradient:
    factor: float = 1.0


@dataclass(frozen=True)
class TensorGradient:
    gradient: Any
    coefficient: Any


@dataclass(frozen=True)
class SymmetricGradient:
    role: str
    factor: float = 1.0


@dataclass(frozen=True)
class Divergence:
    role: str
    factor: float = 1.0
    coefficient: Any = None

    def __mul__(self, other):
        if np.isscalar(other):
            return Divergence(self.role, self.factor * other, self.coefficient)
        if isinstance(other, Divergence) and self.role != other.role:
            if self.coefficient is not None or other.coefficient is not None:
                raise UnsupportedNativeForm("weighted divergence products are not supported")
            return BilinearTerm("divergence", factor=self.factor * other.factor)
        if self.role == "trial" and isinstance(other, TestValue):
            return BilinearTerm("column_divergence", coefficient=self.coefficient, factor=self.factor * other.factor)
        if self.role == "test" and isinstance(other, TrialValue):
            return BilinearTerm("row_divergence", coefficient=self.coefficient, factor=self.factor * other.factor)
        return NotImplemented

    __rmul__ = __mul__


# ------------------------------------------------------------------ terms
@dataclass(frozen=True)
class LinearTerm:
    kind: str
    coefficient: Any
    factor: float = 1.0

    def __neg__(self):
        return LinearTerm(self.kind, self.coefficient, -self.factor)

    def __mul__(self, value):
        if np.isscalar(value):
            return LinearTerm(self.kind, self.coefficient, self.factor * value)
        return NotImplemented

    __rmul__ = __mul__

    def __add__(self, other):
        return LinearSum((self,)) + other

    def __sub__(self, other):
        return self + (-other)


@dataclass(frozen=True)
class LinearSum:
    terms: tuple

    def __add__(self, other):
        if isinstance(other, LinearTerm):
            return LinearSum(self.terms + (other,))
        if isinstance(other, LinearSum):
            return LinearSum(self.terms + other.terms)
        raise UnsupportedNativeForm

    __radd__ = __add__

    def __neg__(self):
        return LinearSum(tuple(-t for t in self.terms))

    def __sub__(self, other):
        return self + (-other)


@dataclass(frozen=True)
class BilinearTerm:
    kind: str
    coefficient: Any = None
    factor: float = 1.0

    def __mul__(self, other):
        if np.isscalar(other):
            return BilinearTerm(self.kind, self.coefficient, self.factor * other)
        if is_symbolic(other):
            return BilinearTerm(self.kind, other, self.factor)
        if isinstance(other, np.ndarray) or (hasattr(other, "value") and hasattr(other, "grad")):
            if self.coefficient is not None:
                raise UnsupportedNativeForm("multiple bilinear coefficients are not supported")
            return BilinearTerm(self.kind, np.asarray(other), self.factor)
        return NotImplemented

    __rmul__ = __mul__

    def __neg__(self):
        return BilinearTerm(self.kind, self.coefficient, -self.factor)

    def __add__(self, other):
        if isinstance(other, BilinearTerm):
            return BilinearSum((self, other))
        if isinstance(other, BilinearSum):
            return BilinearSum((self,) + other.terms)
        return NotImplemented

    def __sub__(self, other):
        return self + (-other)


@dataclass(frozen=True)
class BilinearSum:
    terms: tuple

    def __add__(self, other):
        if isinstance(other, BilinearTerm):
            return BilinearSum(self.terms + (other,))
        if isinstance(other, BilinearSum):
            return BilinearSum(self.terms + other.terms)
        return NotImplemented

    __radd__ = __add__

    def __neg__(self):
        return BilinearSum(tuple(-t for t in self.terms))

    def __sub__(self, other):
        return self + (-other)


# ------------------------------------------------------------------ values
class QuadratureValue:
    """Numerical form value that keeps typed multiplication semantics."""

    __array_priority__ = 1000

    def __init__(self, value, *, parameter_name=None):
        self._value = value
        self.parameter_name = parameter_name

    @property
    def value(self):
        v = self._value
        if callable(v) and not isinstance(v, np.ndarray):
            v = self._value = np.asarray(v())
        elif not isinstance(v, np.ndarray):
            v = self._value = np.asarray(v)
        return v

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.value, dtype=dtype)

    def __getitem__(self, key):
        if self.parameter_name is not None:
            if isinstance(key, bool) or not isinstance(key, (int, np.integer)):
                raise ValueError(f"form parameter {self.parameter_name!r} component index must be an integer")
            try:
                value = self.value[key]
            except IndexError as error:
                raise ValueError(f"form parameter {self.parameter_name!r} has no component {key}") from error
            return QuadratureValue(value)
        return QuadratureValue(self.value[key])

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        values = [i.value if isinstance(i, QuadratureValue) else i for i in inputs]
        return QuadratureValue(getattr(ufunc, method)(*values, **kwargs))

    def __mul__(self, other):
        if isinstance(other, (BilinearTerm, CompositeBilinearTerm)):
            return other * np.asarray(self.value)
        if isinstance(other, CompositeField):
            return CompositeWeightedField(other, np.asarray(self.value))
        if isinstance(other, Divergence):
            return Divergence(other.role, other.factor, np.asarray(self.value))
        return QuadratureValue(self.value * np.asarray(other))

    def __rmul__(self, other):
        if isinstance(other, (BilinearTerm, CompositeBilinearTerm)):
            return other * np.asarray(self.value)
        return QuadratureValue(np.asarray(other) * self.value)

    def __add__(self, other):
        return QuadratureValue(self.value + np.asarray(other))

    def __radd__(self, other):
        return QuadratureValue(np.asarray(other) + self.value)

    def __sub__(self, other):
        return QuadratureValue(self.value - np.asarray(other))

    def __rsub__(self, other):
        return QuadratureValue(np.asarray(other) - self.value)

    def __truediv__(self, other):
        return QuadratureValue(self.value / np.asarray(other))

    def __rtruediv__(self, other):
        return QuadratureValue(np.asarray(other) / self.value)

    def __pow__(self, other):
        return QuadratureValue(self.value ** other)

    def __neg__(self):
        return QuadratureValue(-self.value)


class Parameters:
    """Attribute/item namespace returning symbolic descriptors when absent."""

    def __init__(self, geometry=None):
        self._geometry = {} if geometry is None else geometry

    def __getattr__(self, name):
        if name.startswith("__"):
            raise AttributeError(name)
        if name in self._geometry:
            return self._geometry[name]
        return Coefficient(name)

    def __getitem__(self, name):
        if name in self._geometry:
            return self._geometry[name]
        return Coefficient(name)


def parameter_values(geometry, kwargs):
    values = dict(geometry)
    for name, value in kwargs.items():
        values[name] = (value if callable(value) or (hasattr(value, "value") and hasattr(value, "grad"))
                        else QuadratureValue(value, parameter_name=name))
    return values


def trace_expression(function, fields, geometry, parameters, *, context):
    try:
        return function(*fields, Parameters(parameter_values(geometry, parameters)))
    except (UnsupportedNativeForm, ValueError):
        raise
    except Exception as error:
        raise UnsupportedNativeForm(f"{context} contains an unsupported operation: {error}") from error


def extract_terms(expression, term_type, sum_type, *, message):
    if isinstance(expression, sum_type):
        return expression.terms
    if isinstance(expression, term_type):
        return (expression,)
    raise UnsupportedNativeForm(message)


# ------------------------------------------------------------------ composite fields
# Subfields of a composite basis (python/skfemntv/_composite_fields.py,
# _form_terms.py:117-196): a field is (role, index, kind); products of a
# trial and a test field give a block term (row = test field, column = trial
# field) of kind value / gradient / row_divergence / column_divergence.
def _block(left, right, kind):
    if not isinstance(left, CompositeField) or not isinstance(right, CompositeField):
        raise UnsupportedNativeForm("composite contraction requires two subfields")
    if left.role == right.role:
        raise UnsupportedNativeForm("composite contraction requires trial and test subfields")
    test, trial = (left, right) if left.role == "test" else (right, left)
    return CompositeBilinearTerm(test.field, trial.field, kind)


def _divergence_block(value, divergence):
    field = divergence.field
    if value.role == field.role:
        raise UnsupportedNativeForm("divergence coupling requires trial and test subfields")
    if field.role == "test":
        return CompositeBilinearTerm(field.field, value.field, "row_divergence")
    return CompositeBilinearTerm(value.field, field.field, "column_divergence")


@dataclass(frozen=True)
class CompositeField:
    role: str
    field: int
    kind: str = "value"

    def __mul__(self, other):
        if isinstance(other, CompositeDivergence):
            return _divergence_block(self, other)
        if isinstance(other, CompositeField):
            if self.kind != "value" or other.kind != "value":
                return NotImplemented
            return _block(self, other, "value")
        if is_symbolic(other):
            return CompositeWeightedField(self, other)
        if np.isscalar(other) or hasattr(other, "__array__") or (hasattr(other, "value") and hasattr(other, "grad")):
            return CompositeWeightedField(self, np.asarray(other))
        return NotImplemented

    __rmul__ = __mul__

    def __neg__(self):
        return CompositeWeightedField(self, -1.0)


@dataclass(frozen=True)
class CompositeWeightedField:
    field: CompositeField
    coefficient: Any

    def __mul__(self, other):
        if isinstance(other, CompositeDivergence):
            term = _divergence_block(self.field, other)
        elif isinstance(other, CompositeField):
            term = _block(self.field, other, "value")
        else:
            return NotImplemented
        return CompositeBilinearTerm(term.row_field, term.column_field, term.kind, self.coefficient, term.factor)

    __rmul__ = __mul__

    def _linear_term(self):
        if self.field.role != "test":
            raise UnsupportedNativeForm("composite LinearForm requires test subfields")
        return CompositeLinearTerm(self.field.field, self.field.kind, self.coefficient)

    def __add__(self, other):
        return self._linear_term() + other

    __radd__ = __add__

    def __neg__(self):
        if isinstance(self.coefficient, str) or is_symbolic(self.coefficient):
            return CompositeLinearTerm(self.field.field, self.field.kind, self.coefficient, -1.0)
        return CompositeWeightedField(self.field, -np.asarray(self.coefficient))

    def __sub__(self, other):
        return self + (-other)


@dataclass(frozen=True)
class CompositeDivergence:
    field: CompositeField

    def __mul__(self, other):
        if isinstance(other, (CompositeField, CompositeWeightedField)):
            return other * self
        return NotImplemented

    __rmul__ = __mul__


@dataclass(frozen=True)
class CompositeBilinearTerm:
    row_field: int
    column_field: int
    kind: str
    coefficient: Any = None
    factor: float = 1.0

    def __mul__(self, other):
        if np.isscalar(other):
            return CompositeBilinearTerm(self.row_field, self.column_field, self.kind, self.coefficient,
                                         self.factor * other)
        if is_symbolic(other):
            return CompositeBilinearTerm(self.row_field, self.column_field, self.kind, other, self.factor)
        if isinstance(other, np.ndarray) or (hasattr(other, "value") and hasattr(other, "grad")):
            if self.coefficient is not None:
                raise UnsupportedNativeForm("multiple composite coefficients are not supported")
            return CompositeBilinearTerm(self.row_field, self.column_field, self.kind, np.asarray(other),
                                         self.factor)
        return NotImplemented

    __rmul__ = __mul__

    def __neg__(self):
        return CompositeBilinearTerm(self.row_field, self.column_field, self.kind, self.coefficient, -self.factor)

    def __add__(self, other):
        if isinstance(other, CompositeBilinearTerm):
            return CompositeBilinearSum((self, other))
        if isinstance(other, CompositeBilinearSum):
            return CompositeBilinearSum((self,) + other.terms)
        return NotImplemented

    def __sub__(self, other):
        return self + (-other)


@dataclass(frozen=True)
class CompositeBilinearSum:
    terms: tuple

    def __add__(self, other):
        if isinstance(other, CompositeBilinearTerm):
            return CompositeBilinearSum(self.terms + (other,))
        if isinstance(other, CompositeBilinearSum):
            return CompositeBilinearSum(self.terms + other.terms)
        return NotImplemented

    __radd__ = __add__

    def __neg__(self):
        return CompositeBilinearSum(tuple(-t for t in self.terms))

    def __sub__(self, other):
        return self + (-other)


@dataclass(frozen=True)
class CompositeLinearTerm:
    field: int
    kind: str
    coefficient: Any
    factor: float = 1.0

    def __mul__(self, other):
        if np.isscalar(other):
            return CompositeLinearTerm(self.field, self.kind, self.coefficient, self.factor * other)
        return NotImplemented

    __rmul__ = __mul__

    def __neg__(self):
        return CompositeLinearTerm(self.field, self.kind, self.coefficient, -self.factor)

    def __add__(self, other):
        if isinstance(other, CompositeWeightedField):
            other = other._linear_term()
        if isinstance(other, CompositeLinearTerm):
            return CompositeLinearSum((self, other))
        if isinstance(other, CompositeLinearSum):
            return CompositeLinearSum((self,) + other.terms)
        return NotImplemented

    __radd__ = __add__

    def __sub__(self, other):
        return self + (-other)


@dataclass(frozen=True)
class CompositeLinearSum:
    terms: tuple

    def __add__(self, other):
        if isinstance(other, CompositeWeightedField):
            other = other._linear_term()
        if isinstance(other, CompositeLinearTerm):
            return CompositeLinearSum(self.terms + (other,))
        if isinstance(other, CompositeLinearSum):
            return CompositeLinearSum(self.terms + other.terms)
        return NotImplemented

    __radd__ = __add__

    def __neg__(self):
        return CompositeLinearSum(tuple(-t for t in self.terms))

    def __sub__(self, other):
        return self + (-other)
