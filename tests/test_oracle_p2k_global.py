"""CPU pin of the oracle's global P2 variable-coefficient apply (P2ElementwiseOperator with p2_div_k_grad_affine_q4,
elementwiseoperators/P2ElementwiseOperator.cpp:100-310) against an independent global assembly.

tests/bruteforce.py builds the refined micro-mesh element by element (micro-vertices identified by position, not
by the oracle's indexing), evaluates the coefficient at the physical quadrature points of every micro-element and
assembles the global matrix with scipy. A coefficient that varies in space makes every error of the per-element
anchor, type or node mapping visible (a translated or mis-typed micro-element sees a different k):
  * kind 1, k = 1 + x/2 + y^2/4: the integrand k grad phi_i . grad phi_j has degree 4, so any rule exact to degree
    4 gives the same matrix; the brute force uses an independent collapsed Gauss rule (exact to degree 7);
  * kind 0 without the per-element factor, k = k0 exp(alpha (1 - |x|)): not a polynomial, so the brute force uses
    the oracle's 11-point rule (its table is pinned by the kind-1 case) with its own geometry and assembly.
Meshes: the 24-tet cube, a tilted pyramid and the spherical shell (3, 2) (curved macro-vertices), P2 levels 1-2.
"""
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

import bruteforce as bf
import oracle_lib as ol

MESHDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "meshes")
SHELL = ("shell", 3, 2, 0.55, 1.0)


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


def _model(name, om, L):
    if name == "shell":
        xyz, _ = om.coords(0)
        nodes = {i: xyz[i] for i in range(len(xyz))}
        tets = [list(c) for c in om.cells()]
        return bf.P2Model((nodes, tets), L, scale_key="tree")
    return bf.P2Model(os.path.join(MESHDIR, name), L)


def _oracle_mesh(name):
    return ol.Mesh(*SHELL) if name == "shell" else ol.Mesh(name)


def _kfun(kind):
    if kind == "poly":
        return lambda x: 1.0 + 0.5 * x[:, 0] + 0.25 * x[:, 1] ** 2
    return lambda x: 1.7 * np.exp(2.99 * (1.0 - np.linalg.norm(x, axis=1)))


@pytest.mark.parametrize("name", ["cube_24el.msh", "pyramid_tilted_4el.msh", "shell"])
@pytest.mark.parametrize("L", [1, 2])
@pytest.mark.parametrize("kind", ["poly", "exp"])
def test_variable_coefficient_apply_matches_global_assembly(name, L, kind):
    om = _oracle_mesh(name)
    lat = L + 1
    bm = _model(name, om, L)
    xyz, bnd = om.coords(lat)
    d, idx = cKDTree(bm.lat.xyz).query(xyz)
    assert d.max() < 1e-9 and len(np.unique(idx)) == len(idx) == bm.n
    assert np.array_equal(bnd, bm.boundary[idx])
    rule = bf.tet_gauss_rule(5) if kind == "poly" else ol.p2_quadrature()
    A = bm.stiffness(_kfun(kind), rule)
    if kind == "poly":
        op = ol.P2DivKGradOperator(om, L, L, kind=1)
    else:
        op = ol.P2DivKGradOperator(om, L, L, kind=0, k0=1.7, alpha=2.99, beta=0.0, seed=5)
    n = om.ndofs(lat)
    src, old = np.random.default_rng(3).uniform(-1, 1, n), np.random.default_rng(4).uniform(-1, 1, n)
    s, dst = ol.Function(om, lat, lat), ol.Function(om, lat, lat)
    s.set(lat, src)
    ug = np.zeros(bm.n)
    ug[idx] = src
    y = (A @ ug)[idx]
    for flag, mask in ((ol.INNER, ~bnd), (ol.ALL, np.ones(n, bool))):
        for add in (False, True):
            dst.set(lat, old)
            op.apply(s, dst, lat, flag, add)
            want = np.where(mask, y + (old if add else 0.0), old)
            assert rel_inf(dst.get(lat), want) < 1e-13, (flag, add)


def test_a_translated_coefficient_is_detected():
    """The check above fails for an oracle-like apply whose coefficient is evaluated one micro-cell away: the
    brute force with k(x + h e_x) differs from the true assembly by far more than the tolerance."""
    om = ol.Mesh("cube_24el.msh")
    L = 1
    bm = _model("cube_24el.msh", om, L)
    rule = bf.tet_gauss_rule(5)
    h = 1.0 / (1 << L)
    k = _kfun("poly")
    A = bm.stiffness(k, rule)
    B = bm.stiffness(lambda x: k(x + np.array([h, 0.0, 0.0])), rule)
    u = np.random.default_rng(0).uniform(-1, 1, bm.n)
    assert rel_inf(B @ u, A @ u) > 1e-3
