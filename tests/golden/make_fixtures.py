"""Write golden input/output vectors from the reference implementation.

Runs ONLY in a container that has the reference tree at /root/reference and
its C++ extension built into oracle/_ref (``make -C oracle ref``).  The
reference package is imported under the name ``skfemntv_reference`` with its
own sources and that extension; nothing from it is copied into this
repository -- only the numeric arrays below, saved as ``fixtures.npz``
(numpy, no pickles).

Fixtures (all tiny):
  kuhn_*           MeshTet.init_tensor on a 2x3x2-cell grid (p, t)
  p2_*             MeshTet2.from_mesh of that mesh (p, t)
  dofs_*           Basis(...).element_dofs / N for P1, P2 scalar and P2 vector
  c1_*             P1 Laplace ddot(grad u, grad v) + load dot(source, v) on 2^3
  c2_*             P2 anisotropic dot(mul(A, grad u), grad v), jittered 2^3,
                   per-element SPD A in the recommended (3,3,E,Q) layout, 4-pt rule
  c2n_*            same with a nonsymmetric, x-dependent A in the native (E,Q,3,3) layout
  c3_*             vector P2 2 mu eps:eps + lam div div, jittered 2^3, 4-pt rule
  c4_*             TetN1 maxwell (kappa per element, curl 1) on jittered 2^3:
                   edges, element_edges, signs, CSR, and the mapped values/curls of cell 0
"""

import importlib
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_PY = "/root/reference/python/skfemntv"


def reference_package():
    name = "skfemntv_reference"
    if name in sys.modules:
        return sys.modules[name]
    pkg = types.ModuleType(name)
    pkg.__path__ = [REF_PY, os.path.join(ROOT, "oracle", "_ref")]
    pkg.__package__ = name
    sys.modules[name] = pkg
    return pkg


def main():
    reference_package()
    basis = importlib.import_module("skfemntv_reference.basis")
    forms = importlib.import_module("skfemntv_reference.forms")
    helpers = importlib.import_module("skfemntv_reference.helpers")
    hb = importlib.import_module("skfemntv_reference.hcurl_basis")
    ha = importlib.import_module("skfemntv_reference.hcurl_assembler")
    out = {}

    mesh = basis.MeshTet.init_tensor(np.linspace(0, 1, 3), np.linspace(0, 1, 4), np.linspace(0, 1, 3))
    out["kuhn_p"], out["kuhn_t"] = mesh.p, mesh.t
    m2 = basis.MeshTet2.from_mesh(mesh)
    out["p2_p"], out["p2_t"] = m2.p, m2.t
    for tag, m, el, dim in (("p1", mesh, basis.ElementTetP1(), 1), ("p2", m2, basis.ElementTetP2(), 1),
                            ("vp2", m2, basis.ElementTetP2(), 3)):
        b = basis.Basis(m, basis.ElementVector(el, dim=dim))
        out[f"dofs_{tag}"] = b.element_dofs
        out[f"dofs_{tag}_N"] = np.array(b.N)

    # C1: P1 Laplace + load
    ax = np.linspace(0, 1, 3)
    m = basis.MeshTet.init_tensor(ax, ax, ax)
    b = basis.Basis(m, basis.ElementVector(basis.ElementTetP1(), dim=1))
    lap = forms.BilinearForm(lambda u, v, w: helpers.ddot(helpers.grad(u), helpers.grad(v)))
    load = forms.LinearForm(lambda v, w: helpers.dot(w.source, v))
    K = forms.asm(lap, b).tocsr()
    out["c1_p"], out["c1_t"] = m.p, m.t
    out["c1_indptr"], out["c1_indices"], out["c1_data"] = K.indptr.astype(np.int64), K.indices.astype(np.int64), K.data
    out["c1_rhs"] = np.asarray(forms.asm(load, b, source=np.array([1.])))

    # jittered 2^3 cube shared by C2/C3/C4
    rng = np.random.default_rng(2024)
    p = m.p.copy()
    interior = np.all((p > 1e-12) & (p < 1 - 1e-12), axis=0)
    p[:, interior] += rng.uniform(-0.1, 0.1, size=(3, int(interior.sum())))
    mj = basis.MeshTet(p, m.t)
    m2j = basis.MeshTet2.from_mesh(mj)
    X = np.array([[.5854101966249685, .1381966011250105, .1381966011250105, .1381966011250105],
                  [.1381966011250105, .5854101966249685, .1381966011250105, .1381966011250105],
                  [.1381966011250105, .1381966011250105, .5854101966249685, .1381966011250105]])
    W = np.full(4, 1. / 24.)
    E = mj.nelements
    out["jit_p"], out["jit_t"] = mj.p, mj.t

    # C2: P2 anisotropic, SPD per element, recommended layout
    b2 = basis.Basis(m2j, basis.ElementVector(basis.ElementTetP2(), dim=1), quadrature=(X, W))
    g = rng.standard_normal((E, 3, 3))
    A = np.einsum("eij,ekj->eik", g, g) + 0.5 * np.eye(3)
    Arec = np.ascontiguousarray(np.broadcast_to(np.moveaxis(A, 0, -1)[..., None], (3, 3, E, 4)))
    an = forms.BilinearForm(lambda u, v, w: helpers.dot(helpers.mul(w.diffusion, helpers.grad(u)), helpers.grad(v)))
    K = forms.asm(an, b2, diffusion=Arec).tocsr()
    out["c2_A"] = A
    out["c2_indptr"], out["c2_indices"], out["c2_data"] = K.indptr.astype(np.int64), K.indices.astype(np.int64), K.data
    # C2n: nonsymmetric x-dependent tensor, native layout
    x = np.moveaxis(b2.global_coordinates, -1, 0)
    T = np.empty((E, 4, 3, 3))
    T[..., 0, 0] = 1. + x[0]; T[..., 0, 1] = 0.3 + 0.1 * x[1]; T[..., 0, 2] = 0.
    T[..., 1, 0] = -0.2; T[..., 1, 1] = 2. - 0.4 * x[0]; T[..., 1, 2] = 0.1 * x[2]
    T[..., 2, 0] = 0.05; T[..., 2, 1] = 0.; T[..., 2, 2] = 1.5
    K = forms.asm(an, b2, diffusion=T).tocsr()
    out["c2n_x"] = b2.global_coordinates
    out["c2n_indptr"], out["c2n_indices"], out["c2n_data"] = K.indptr.astype(np.int64), K.indices.astype(np.int64), K.data

    # C3: vector P2 elasticity
    b3 = basis.Basis(m2j, basis.ElementVector(basis.ElementTetP2(), dim=3), quadrature=(X, W))
    Ey = np.exp(rng.standard_normal(E))
    nu = rng.uniform(0.2, 0.45, E)
    lam = Ey * nu / ((1 + nu) * (1 - 2 * nu))
    mu = Ey / (2 * (1 + nu))
    el = forms.BilinearForm(lambda u, v, w: 2. * w.mu * helpers.ddot(helpers.sym_grad(u), helpers.sym_grad(v))
                            + w.lam * (helpers.div(u) * helpers.div(v)))
    K = forms.asm(el, b3, mu=mu[:, None], lam=lam[:, None]).tocsr()
    out["c3_mu"], out["c3_lam"] = mu, lam
    out["c3_indptr"], out["c3_indices"], out["c3_data"] = K.indptr.astype(np.int64), K.indices.astype(np.int64), K.data

    # C4: TetN1 maxwell
    nb = hb.AffineTetN1Basis(mj)
    kappa = rng.uniform(0.1, 10., E)
    K = ha.TetN1Assembler(nb).assemble_maxwell(mass_coefficient=kappa[:, None], curl_coefficient=1.).tocsr()
    out["c4_kappa"] = kappa
    out["c4_edges"] = nb.dof_map.topology.edges
    out["c4_element_edges"] = nb.dof_map.topology.element_edges
    out["c4_signs"] = nb.dof_map.topology.element_edge_signs.astype(np.int8)
    out["c4_indptr"], out["c4_indices"], out["c4_data"] = K.indptr.astype(np.int64), K.indices.astype(np.int64), K.data
    out["c4_values0"] = nb._element_values[0]
    out["c4_curls0"] = nb._element_curls[0]
    out["c4_X"], out["c4_W"] = nb.X, nb.W

    path = os.path.join(HERE, "fixtures.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
