"""CPU oracle for the MI355X element-assembly path -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product: only ``tests/``, the smoke
check in ``__graft_entry__`` and the CPU-baseline leg of ``bench.py`` import
it.  It has three layers:

* ``host``       numpy restatement of the reference's Python-side setup
                 (meshes, P2 numbering, DOFs, quadrature, tabulation, N1 map);
* ``liboracle``  C restatement (``fem_oracle.c``) of the reference's C++
                 kernels (geometry, CSR pattern + map, element kernels,
                 edge topology, N1 integration, scatter, colouring);
* ``_ref``       the reference's own C++ extension, compiled from its sources
                 by ``oracle/Makefile`` where the reference tree exists (this
                 build container only; it never travels to the GPU host).
                 Used to pin the C restatement.

Parity status of each function is listed in ``PARITY`` below.
"""

from __future__ import annotations

import ctypes
import glob
import os
import sys

import numpy as np

from . import host

HERE = os.path.dirname(os.path.abspath(__file__))

# function -> how it is pinned (checked by tests/test_oracle_*.py)
PARITY = {
    "host.kuhn_mesh": "golden fixture from the reference + volume/conformity invariants",
    "host.p2_mesh": "golden fixture from the reference + first-seen/midpoint invariants",
    "host.element_dofs": "golden fixture from the reference + interleaving invariant",
    "host.tet_rule_4/keast11/duffy": "monomial exactness closed forms",
    "host.tabulate_p1/p2": "partition of unity, Kronecker property at nodes, gradient finite differences",
    "host.n1_reference": "edge-moment duality (identity) closed form",
    "host.n1_basis": "golden fixture from the reference + reference edge moments + constant-field reproduction",
    "host.anisotropic_tensor": "golden fixture (layout) from the reference",
    "c.tabulate_geometry": "reference C++ (_ref) on the same inputs, bit-exact",
    "c.pattern_build": "reference C++ (_ref) on the same inputs, bit-exact",
    "c.assemble_bilinear": "reference C++ (_ref) on the same inputs + closed forms",
    "c.assemble_tensor": "reference C++ (_ref) on the same inputs + closed forms",
    "c.assemble_cross": "reference C++ (_ref) CrossBilinearAssembler on the same inputs + closed forms",
    "c.assemble_linear": "reference C++ (_ref) on the same inputs + closed forms",
    "c.edge_topology": "reference C++ (_ref) on the same inputs, bit-exact",
    "c.integrate_tet_n1": "reference C++ (_ref) on the same inputs",
    "c.scatter_add": "numpy.add.at on the same inputs",
    "cut.cut_pattern": "reference C++ (_ref) CutBilinearFormAssembler/CutCrossAssembler, bit-exact",
    "cut.cut_bilinear": "reference C++ (_ref) on the same inputs + partition of unity, zero stiffness row sums",
    "cut.cut_linear": "reference C++ (_ref) on the same inputs + partition of unity, K x = load of grad x",
    "cut.cut_cross": "reference C++ (_ref) CutCrossAssembler on the same inputs (scalar and tensor)",
}

_lib = None


def build():
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir("/root/reference/cpp/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        _lib = ctypes.CDLL(path)
        _lib.oracle_last_error.restype = ctypes.c_char_p
        _lib.oracle_colouring.restype = ctypes.c_int64
    return _lib


def ref_module():
    """Import the reference's compiled `_skfn` from oracle/_ref (or None)."""
    cands = glob.glob(os.path.join(HERE, "_ref", "_skfn*.so"))
    if not cands:
        return None
    d = os.path.join(HERE, "_ref")
    if d not in sys.path:
        sys.path.insert(0, d)
    import _skfn  # noqa: E402
    return _skfn


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise ValueError(lib().oracle_last_error().decode())


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def tabulate_geometry(p, t, shape, refgrad, gshape, ggrad, W):
    """C restatement of tabulate_basis_geometry; returns the same 7-tuple plus
    (bad_cell, bad_qp, bad_reason)."""
    p, t = _f64(p), _i64(t)
    shape, refgrad, gshape, ggrad, W = map(_f64, (shape, refgrad, gshape, ggrad, W))
    d, P = p.shape
    gn, E = t.shape
    Q, n = shape.shape
    o_shape = np.zeros((E, Q, n)); o_grad = np.zeros((E, Q, n, d)); o_dx = np.zeros((E, Q))
    o_x = np.zeros((E, Q, d)); o_det = np.zeros((E, Q)); o_tol = np.zeros((E, Q)); o_cond = np.zeros((E, Q))
    bc, bq, br = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    L = lib()
    L.oracle_tabulate_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p] + [ctypes.c_void_p] * 7 + [ctypes.c_void_p] * 3
    _check(L.oracle_tabulate_geometry(_p(p), P, d, _p(t), gn, E, _p(shape), _p(refgrad), Q, n, _p(gshape),
                                      _p(ggrad), _p(W), _p(o_shape), _p(o_grad), _p(o_dx), _p(o_x), _p(o_det),
                                      _p(o_tol), _p(o_cond), ctypes.byref(bc), ctypes.byref(bq), ctypes.byref(br)))
    return (o_shape, o_grad, o_dx, o_x, o_det, o_tol, o_cond), (bc.value, bq.value, br.value)


def pattern(rows, cols=None, nrows=None):
    """CSR pattern + element->CSR map.  rows/cols (E,L) entity-major."""
    rows = _i64(rows)
    cols = rows if cols is None else _i64(cols)
    E, Lr = rows.shape
    Lc = cols.shape[1]
    if nrows is None:
        nrows = int(rows.max(initial=-1)) + 1
    ip, ix, sc = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nnz = ctypes.c_int64()
    L = lib()
    L.oracle_pattern_build.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
    _check(L.oracle_pattern_build(_p(rows), _p(cols), E, Lr, Lc, nrows, ctypes.byref(ip), ctypes.byref(ix),
                                  ctypes.byref(sc), ctypes.byref(nnz)))

    def take(ptr, count):
        arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int64)), shape=(count,)).copy()
        L.oracle_free(ptr)
        return arr

    n = nnz.value
    return take(ip, nrows + 1), take(ix, max(n, 1))[:n], take(sc, max(E * Lr * Lc, 1))[:E * Lr * Lc].reshape(E, Lr, Lc)


def colouring(rows, nrows):
    rows = _i64(rows)
    E, Lr = rows.shape
    colour = np.zeros(E, dtype=np.int32)
    L = lib()
    L.oracle_colouring.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
    ncol = L.oracle_colouring(_p(rows), E, Lr, nrows, _p(colour))
    return colour, int(ncol)


def assemble_bilinear(shape, grads, dx, scatter, nnz, C, value=None, gradient=None, symgrad=None,
                      divergence=None, threads=1, colour=None):
    shape, grads, dx = map(_f64, (shape, grads, dx))
    value, gradient, symgrad, divergence = map(_f64, (value, gradient, symgrad, divergence))
    scatter = _i64(scatter)
    E, Q, n = shape.shape
    d = grads.shape[3]
    out = np.zeros(nnz)
    ncol = 0
    if colour is not None:
        colour = np.ascontiguousarray(colour, dtype=np.int32)
        ncol = int(colour.max(initial=-1)) + 1
    L = lib()
    L.oracle_assemble_bilinear.argtypes = [ctypes.c_void_p] * 8 + [ctypes.c_int64] * 5 + [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
    _check(L.oracle_assemble_bilinear(_p(shape), _p(grads), _p(dx), _p(value), _p(gradient), _p(symgrad),
                                      _p(divergence), _p(scatter), E, n, C, Q, d, _p(out), nnz,
                                      _p(colour), ncol, int(threads)))
    return out


def assemble_tensor(grads, dx, tensor, scatter, nnz, threads=1, colour=None):
    grads, dx, tensor = map(_f64, (grads, dx, tensor))
    scatter = _i64(scatter)
    E, Q, n, d = grads.shape
    out = np.zeros(nnz)
    ncol = 0
    if colour is not None:
        colour = np.ascontiguousarray(colour, dtype=np.int32)
        ncol = int(colour.max(initial=-1)) + 1
    L = lib()
    L.oracle_assemble_tensor.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64] * 4 + [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
    _check(L.oracle_assemble_tensor(_p(grads), _p(dx), _p(tensor), _p(scatter), E, n, Q, d, _p(out), nnz,
                                    _p(colour), ncol, int(threads)))
    return out


def assemble_cross(rshape, rgrads, cshape, cgrads, dx, scatter, nnz, Cr, Cc, row_kind, col_kind,
                   coefficient=None, tensor=False):
    """General cross kernel (serial).  kinds 'value'/'gradient'; coefficient
    None, scalar (E,Q), or tensor (E,Q,Cr,[d],Cc,[d]) when tensor=True."""
    rshape, rgrads, cshape, cgrads, dx = map(_f64, (rshape, rgrads, cshape, cgrads, dx))
    coefficient = _f64(coefficient)
    scatter = _i64(scatter)
    E, Q, nr = rshape.shape
    nc = cshape.shape[2]
    d = rgrads.shape[3]
    out = np.zeros(nnz)
    L = lib()
    L.oracle_assemble_cross.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int] * 3 + [ctypes.c_void_p] + [
        ctypes.c_int64] * 7 + [ctypes.c_void_p, ctypes.c_int64]
    _check(L.oracle_assemble_cross(_p(rshape), _p(rgrads), _p(cshape), _p(cgrads), _p(dx), _p(coefficient),
                                   int(bool(tensor)), int(row_kind == "gradient"), int(col_kind == "gradient"),
                                   _p(scatter), E, nr, Cr, nc, Cc, Q, d, _p(out), nnz))
    return out


def assemble_linear(shape, grads, dx, dofs, ndofs, C, value=None, gradient=None):
    """dofs (E,n,C).  Constant coefficients are (C,) / (C,d)."""
    shape, grads, dx = map(_f64, (shape, grads, dx))
    dofs = _i64(dofs)
    E, Q, n = shape.shape
    d = grads.shape[3]
    cv = value is not None and np.ndim(value) == 1
    cg = gradient is not None and np.ndim(gradient) == 2
    value, gradient = _f64(value), _f64(gradient)
    out = np.zeros(ndofs)
    L = lib()
    L.oracle_assemble_linear.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64] * 5 + [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64]
    _check(L.oracle_assemble_linear(_p(shape), _p(grads), _p(dx), _p(dofs), E, n, C, Q, d, _p(value), int(cv),
                                    _p(gradient), int(cg), _p(out), ndofs))
    return out


def edge_topology(t4):
    """(edges (2,nE), element_edges (6,E), signs (6,E) int8)."""
    t4 = _i64(t4)
    E = t4.shape[1]
    edges = np.zeros((2, 6 * E), dtype=np.int64)
    ee = np.zeros((6, E), dtype=np.int64)
    signs = np.zeros((6, E), dtype=np.int8)
    nE = ctypes.c_int64()
    L = lib()
    L.oracle_edge_topology.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
    _check(L.oracle_edge_topology(_p(t4), E, _p(edges), _p(ee), _p(signs), ctypes.byref(nE)))
    return edges[:, :nE.value].copy(), ee, signs


def integrate_tet_n1(values, curls, dx, mass, curl, include_mass=True, include_curl=True):
    values, curls, dx, mass, curl = map(_f64, (values, curls, dx, mass, curl))
    E, _, _, Q = values.shape
    out = np.zeros((E, 6, 6))
    L = lib()
    L.oracle_integrate_tet_n1.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                                                  ctypes.c_int64, ctypes.c_void_p]
    _check(L.oracle_integrate_tet_n1(_p(values), _p(curls), _p(dx), _p(mass), _p(curl), int(include_mass),
                                     int(include_curl), E, Q, _p(out)))
    return out


def scatter_add(scatter, elements, nnz):
    scatter, elements = _i64(scatter).ravel(), _f64(elements).ravel()
    data = np.zeros(nnz)
    L = lib()
    L.oracle_scatter_add.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                     ctypes.c_int64]
    _check(L.oracle_scatter_add(_p(scatter), _p(elements), scatter.size, _p(data), nnz))
    return data
