"""Symbolic tracing of user forms.

``asm`` calls the user's form function once, with symbolic stand-ins for the
trial and test functions (or for the subfields of a composite basis) and a
parameter namespace ``w``.  What the function returns is an expression of
three kinds of objects:

``Field``
    one occurrence of a trial or test function: its role, the composite
    subfield it belongs to (``None`` for a plain basis), the operator applied
    to it (value, gradient, symmetric gradient, divergence, or a gradient
    multiplied by a tensor), a numeric scale and an optional weight.
``Term``
    a contraction the assembler implements, labelled with the reference's
    term kinds (value, gradient, symmetric_gradient, divergence,
    row_divergence, column_divergence, gradient_tensor); it carries the
    coefficient, the scale and, on composite bases, the test and trial
    subfields.  A term of a linear form has a test side only.
``TermSum``
    a sum of terms.

Numbers from the parameters and the geometry enter as ``PointValue`` (an
array evaluated when first needed, with numpy arithmetic), parameters that
were not supplied as ``Param`` placeholders that are resolved when the term
is assembled.  Which contractions exist and which messages reject the
others follow the reference's typed-form rules (python/skfemntv/
_h1_fields.py, _form_terms.py, _composite_fields.py, forms.py:788-1091);
the classes here are this package's own.
"""

from __future__ import annotations

from typing import Any, Mapping

import numpy as np


class UnsupportedNativeForm(Exception):
    """Raised when a form cannot be assembled by the native backend."""


VALUE, GRAD, SYMGRAD, DIV, TGRAD = "value", "grad", "symgrad", "div", "tgrad"
TRIAL, TEST = "trial", "test"


def _number(x):
    return isinstance(x, (int, float, np.integer, np.floating)) and not isinstance(x, bool)


# ------------------------------------------------------------------ values
class Param:
    """Placeholder of a form parameter that was not supplied (optionally one
    component of it); asm resolves it against the keyword arguments."""

    __slots__ = ("name", "index")
    __array_priority__ = 1000

    def __init__(self, name, index=None):
        self.name, self.index = name, index

    def __getitem__(self, index):
        if isinstance(index, bool) or not isinstance(index, (int, np.integer)):
            raise ValueError("coefficient component index must be an integer")
        return Param(self.name, int(index))

    def __mul__(self, other):
        if isinstance(other, (Field, Term, TermSum)):
            return other.weighted(self)
        return NotImplemented

    __rmul__ = __mul__

    def __repr__(self):
        return f"Param({self.name!r})" if self.index is None else f"Param({self.name!r}, {self.index})"


class PointValue:
    """A numeric form value (parameter or geometry) at the quadrature points.
    Arithmetic stays numeric; products with fields and terms weight them."""

    __array_priority__ = 1000

    def __init__(self, value, name=None):
        self._data = value
        self.name = name

    @property
    def value(self):
        data = self._data
        if callable(data) and not isinstance(data, np.ndarray):
            data = np.asarray(data())
        elif not isinstance(data, np.ndarray):
            data = np.asarray(data)
        self._data = data
        return data

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.value, dtype=dtype)

    def __getitem__(self, key):
        if self.name is not None:
            if isinstance(key, bool) or not isinstance(key, (int, np.integer)):
                raise ValueError(f"form parameter {self.name!r} component index must be an integer")
            try:
                return PointValue(self.value[key])
            except IndexError as error:
                raise ValueError(f"form parameter {self.name!r} has no component {key}") from error
        return PointValue(self.value[key])

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        args = [x.value if isinstance(x, PointValue) else x for x in inputs]
        return PointValue(getattr(ufunc, method)(*args, **kwargs))

    def _numeric(self, other, fn):
        if isinstance(other, (Field, Term, TermSum)):
            return other.weighted(self.value)
        return PointValue(fn(self.value, np.asarray(other)))

    def __mul__(self, other):
        return self._numeric(other, lambda a, b: a * b)

    def __rmul__(self, other):
        return self._numeric(other, lambda a, b: b * a)

    def __add__(self, other):
        return PointValue(self.value + np.asarray(other))

    def __radd__(self, other):
        return PointValue(np.asarray(other) + self.value)

    def __sub__(self, other):
        return PointValue(self.value - np.asarray(other))

    def __rsub__(self, other):
        return PointValue(np.asarray(other) - self.value)

    def __truediv__(self, other):
        return PointValue(self.value / np.asarray(other))

    def __rtruediv__(self, other):
        return PointValue(np.asarray(other) / self.value)

    def __pow__(self, other):
        return PointValue(self.value ** other)

    def __neg__(self):
        return PointValue(-self.value)


def _is_discrete_field(x):
    # a finite-element function evaluated at the points (value and gradient)
    return hasattr(x, "value") and hasattr(x, "grad")


def coefficient_of(x):
    """The coefficient a value contributes to a term: a Param stays
    symbolic, numbers and arrays become float arrays; None if `x` is not a
    coefficient."""
    if isinstance(x, Param):
        return x
    if isinstance(x, PointValue):
        return np.asarray(x.value)
    if _number(x) or isinstance(x, np.ndarray) or _is_discrete_field(x):
        return np.asarray(x)
    return None


# ------------------------------------------------------------------ fields
class Field:
    """One occurrence of a trial or test function in a form."""

    __slots__ = ("role", "sub", "op", "scale", "weight")

    def __init__(self, role, sub=None, op=VALUE, scale=1.0, weight=None):
        self.role, self.sub, self.op, self.scale, self.weight = role, sub, op, scale, weight

    def _with(self, **changes):
        new = Field(self.role, self.sub, self.op, self.scale, self.weight)
        for key, value in changes.items():
            setattr(new, key, value)
        return new

    @property
    def composite(self):
        return self.sub is not None

    def weighted(self, coef):
        if self.weight is not None:
            raise UnsupportedNativeForm("multiple composite coefficients are not supported" if self.composite
                                        else "multiple bilinear coefficients are not supported")
        return self._with(weight=coef)

    def __mul__(self, other):
        if _number(other):
            return self._with(scale=self.scale * other)
        if isinstance(other, Field):
            return product(self, other)
        if isinstance(other, (Term, TermSum)):
            return NotImplemented
        coef = coefficient_of(other)
        if coef is None:
            return NotImplemented
        if self.composite or self.op == DIV:
            return self.weighted(coef)
        raise UnsupportedNativeForm("a plain trial or test field is weighted through dot() or a contraction")

    __rmul__ = __mul__

    def __neg__(self):
        return self._with(scale=-self.scale)

    # a weighted test subfield on its own is a linear value term
    def as_linear_term(self):
        if not (self.composite and self.op == VALUE):
            raise UnsupportedNativeForm("composite LinearForm must reduce to subfield value or gradient "
                                        "contractions")
        if self.role != TEST:
            raise UnsupportedNativeForm("composite LinearForm requires test subfields")
        return Term(VALUE, self.weight, self.scale, test_sub=self.sub, linear=True)

    def __add__(self, other):
        return self.as_linear_term() + other

    def __radd__(self, other):
        return other + self.as_linear_term()

    def __sub__(self, other):
        return self.as_linear_term() - other


# ------------------------------------------------------------------ terms
class Term:
    """A contraction with a kernel: kind, coefficient, scale, and the test /
    trial subfields on composite bases (None on a plain basis)."""

    __slots__ = ("kind", "coef", "scale", "test_sub", "trial_sub", "linear")

    def __init__(self, kind, coef=None, scale=1.0, test_sub=None, trial_sub=None, linear=False):
        self.kind, self.coef, self.scale = kind, coef, scale
        self.test_sub, self.trial_sub, self.linear = test_sub, trial_sub, linear

    def _with(self, **changes):
        new = Term(self.kind, self.coef, self.scale, self.test_sub, self.trial_sub, self.linear)
        for key, value in changes.items():
            setattr(new, key, value)
        return new

    def weighted(self, coef):
        if self.coef is not None:
            raise UnsupportedNativeForm("multiple composite coefficients are not supported"
                                        if self.test_sub is not None else
                                        "multiple bilinear coefficients are not supported")
        return self._with(coef=coef)

    def __mul__(self, other):
        if _number(other):
            return self._with(scale=self.scale * other)
        if self.linear:
            return NotImplemented
        coef = coefficient_of(other)
        if coef is None:
            return NotImplemented
        return self.weighted(coef)

    __rmul__ = __mul__

    def __neg__(self):
        return self._with(scale=-self.scale)

    def __add__(self, other):
        return TermSum((self,)) + other

    def __radd__(self, other):
        return other + TermSum((self,)) if not isinstance(other, (int, float)) or other != 0 else TermSum((self,))

    def __sub__(self, other):
        return self + (-other)


class TermSum:
    __slots__ = ("terms",)

    def __init__(self, terms):
        self.terms = tuple(terms)

    def __add__(self, other):
        if isinstance(other, (int, float)) and other == 0:
            return self
        if isinstance(other, Field):
            other = other.as_linear_term()
        if isinstance(other, Term):
            other = TermSum((other,))
        if not isinstance(other, TermSum):
            return NotImplemented
        mixed = {t.linear for t in self.terms + other.terms}
        if len(mixed) > 1:
            raise UnsupportedNativeForm("a form cannot add linear and bilinear terms")
        return TermSum(self.terms + other.terms)

    __radd__ = __add__

    def __neg__(self):
        return TermSum(-t for t in self.terms)

    def __sub__(self, other):
        return self + (-other)

    def __mul__(self, other):
        if _number(other):
            return TermSum(t * other for t in self.terms)
        return NotImplemented

    __rmul__ = __mul__

    def weighted(self, coef):
        return TermSum(t.weighted(coef) for t in self.terms)


# ------------------------------------------------------------------ operators
def apply_grad(x):
    if isinstance(x, Field):
        if x.op != VALUE:
            raise UnsupportedNativeForm("grad() requires a form field")
        return x._with(op=GRAD)
    try:
        return x.grad
    except AttributeError as error:
        raise UnsupportedNativeForm("grad() requires a form field") from error


def apply_symgrad(x):
    """Symmetric gradient; numeric gradients give 0.5 (G + G^T)."""
    if isinstance(x, Field):
        if x.composite:
            raise UnsupportedNativeForm("sym_grad() of a composite subfield is not supported")
        if x.op != VALUE:
            raise UnsupportedNativeForm("grad() requires a form field")
        return x._with(op=SYMGRAD)
    g = apply_grad(x)
    return 0.5 * (g + np.einsum("ij...->ji...", g))


def apply_div(x):
    if isinstance(x, Field):
        if x.op != VALUE:
            raise UnsupportedNativeForm("div() expects a composite value" if x.composite
                                        else "div() requires a vector field")
        return x._with(op=DIV)
    try:
        return np.einsum("ii...->...", x.grad)
    except AttributeError as error:
        raise UnsupportedNativeForm("div() requires a vector field") from error


def apply_mul(coef, x):
    """mul(A, grad(u)): a tensor applied to a trial or test gradient."""
    c = None if isinstance(coef, Field) else coefficient_of(coef)
    if isinstance(x, Field) and x.op == GRAD and not x.composite and c is not None and not _number(coef):
        return x._with(op=TGRAD, weight=c)
    raise UnsupportedNativeForm("native mul currently supports coefficient @ trial/test gradient")


def _roles(a, b, message):
    if a.role == b.role:
        raise UnsupportedNativeForm(message)
    return (a, b) if a.role == TEST else (b, a)


def _block(test, trial, kind, coef=None):
    scale = test.scale * trial.scale
    if test.weight is not None and trial.weight is not None:
        raise UnsupportedNativeForm("multiple composite coefficients are not supported")
    if coef is None:
        coef = test.weight if test.weight is not None else trial.weight
    return Term(kind, coef, scale, test_sub=test.sub, trial_sub=trial.sub)


def product(a, b):
    """a * b of two fields: divergence couplings and composite value blocks."""
    if a.composite != b.composite:
        raise UnsupportedNativeForm("plain and composite fields cannot be combined")
    if a.op == DIV and b.op == DIV and not a.composite:
        test, trial = _roles(a, b, "a divergence product requires trial and test fields")
        if a.weight is not None or b.weight is not None:
            raise UnsupportedNativeForm("weighted divergence products are not supported")
        return Term("divergence", None, a.scale * b.scale)
    for d, other in ((a, b), (b, a)):
        if d.op == DIV and other.op == VALUE:
            test, trial = _roles(d, other, "divergence coupling requires trial and test subfields")
            kind = "row_divergence" if d.role == TEST else "column_divergence"
            if d.composite:
                return _block(test, trial, kind)
            if other.weight is not None:
                raise UnsupportedNativeForm("multiple bilinear coefficients are not supported")
            return Term(kind, d.weight, d.scale * other.scale)
    if a.composite and a.op == VALUE and b.op == VALUE:
        test, trial = _roles(a, b, "composite contraction requires trial and test subfields")
        return _block(test, trial, "value")
    raise UnsupportedNativeForm("the product of two form fields is an unsupported operation; use dot() or ddot()")


def _linear(field, other, kind):
    """A coefficient contracted with a test field."""
    coef = coefficient_of(other)
    if coef is None or field.role != TEST or field.weight is not None:
        return None
    return Term(kind, coef, field.scale, test_sub=field.sub, linear=True)


def contract(a, b, how):
    """dot (how = 'dot') or ddot (how = 'ddot') with at least one field."""
    if isinstance(a, Field) and isinstance(b, Field):
        if a.composite != b.composite:
            raise UnsupportedNativeForm("plain and composite fields cannot be combined")
        if a.composite:
            test, trial = _roles(a, b, "composite contraction requires trial and test subfields")
            if how == "ddot":
                if a.op == GRAD and b.op == GRAD:
                    return _block(test, trial, "gradient")
                raise UnsupportedNativeForm("composite ddot requires two gradients")
            if a.op == b.op and a.op in (VALUE, GRAD):
                return _block(test, trial, "value" if a.op == VALUE else "gradient")
            raise UnsupportedNativeForm("composite dot requires two values or two gradients")
        ops = {a.op, b.op}
        if ops == {TGRAD, GRAD} and how == "dot":
            tensor, plain = (a, b) if a.op == TGRAD else (b, a)
            if tensor.role == plain.role:
                raise UnsupportedNativeForm("anisotropic contraction requires trial and test gradients")
            return Term("gradient_tensor", tensor.weight, tensor.scale * plain.scale)
        if a.role == b.role:
            raise UnsupportedNativeForm("a contraction requires a trial and a test field")
        if a.weight is not None or b.weight is not None:
            raise UnsupportedNativeForm("weighted fields are contracted through a coefficient on the term")
        if ops == {VALUE} and how == "dot":
            return Term("value", None, a.scale * b.scale)
        if ops == {GRAD}:   # dot(grad u, grad v) is accepted as well as ddot
            return Term("gradient", None, a.scale * b.scale)
        if ops == {SYMGRAD} and how == "ddot":
            return Term("symmetric_gradient", None, a.scale * b.scale)
        raise UnsupportedNativeForm(f"unsupported {how}() operands in a native form")
    field, other = (a, b) if isinstance(a, Field) else (b, a)
    want = {("dot", VALUE): "value", ("ddot", GRAD): "gradient"}.get((how, field.op))
    term = _linear(field, other, want) if want else None
    if term is None:
        if field.composite:
            raise UnsupportedNativeForm(f"composite {how} requires a coefficient and a test subfield value"
                                        if how == "dot" else "composite ddot requires two gradients")
        raise UnsupportedNativeForm(f"unsupported {how}() operands in a native form")
    return term


# ------------------------------------------------------------------ tracing
class FormParameters:
    """The ``w`` of a form: geometry values by name, supplied parameters as
    PointValue, anything else as a Param placeholder."""

    def __init__(self, values):
        self._values = values

    def __getattr__(self, name):
        if name.startswith("__"):
            raise AttributeError(name)
        return self[name]

    def __getitem__(self, name):
        values = self.__dict__["_values"]
        return values[name] if name in values else Param(name)


def namespace(geometry, kwargs):
    values = dict(geometry)
    for name, value in kwargs.items():
        values[name] = value if callable(value) or _is_discrete_field(value) else PointValue(value, name=name)
    return FormParameters(values)


def trace(function, args, geometry, kwargs, *, context):
    """Call the form function on symbolic arguments."""
    try:
        return function(*args, namespace(geometry, kwargs))
    except (UnsupportedNativeForm, ValueError):
        raise
    except Exception as error:
        raise UnsupportedNativeForm(f"{context} contains an unsupported operation: {error}") from error


def collect(expression, *, linear, message):
    """The terms of a traced expression (a Term, a TermSum or, for composite
    linear forms, a weighted test subfield)."""
    if linear and isinstance(expression, Field) and expression.composite:
        expression = expression.as_linear_term()
    if isinstance(expression, Term):
        expression = TermSum((expression,))
    if not isinstance(expression, TermSum) or any(t.linear != linear for t in expression.terms):
        raise UnsupportedNativeForm(message)
    return expression.terms


def resolve(value: Any, parameters: Mapping[str, Any]) -> Any:
    """Replace a Param placeholder by the supplied parameter (component)."""
    if isinstance(value, str):
        value = Param(value)
    if not isinstance(value, Param):
        return value
    if value.name not in parameters:
        raise ValueError(f"missing form parameter {value.name!r}")
    supplied = parameters[value.name]
    if value.index is None:
        return supplied
    try:
        return np.asarray(supplied)[value.index]
    except IndexError as error:
        raise ValueError(f"form parameter {value.name!r} has no component {value.index}") from error


def coefficient_array(term, parameters, *, squeeze=False):
    """The term's coefficient times its scale as a float64 array (the bare
    scale when the term has no coefficient); with squeeze, arrays of more
    than two dimensions lose their unit axes."""
    if term.coef is None:
        return term.scale
    arr = np.asarray(resolve(term.coef, parameters), dtype=np.float64)
    if term.scale != 1.0:   # x * 1.0 == x exactly; broadcast views stay views
        arr = term.scale * arr
    if squeeze and arr.ndim > 2:
        arr = np.squeeze(arr)
    return arr
