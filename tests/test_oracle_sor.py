"""CPU checks of the multi-colour SOR ordering (the restatement in sor_model.py): the colouring is proper on the
micro-mesh assembled independently (bruteforce.py), the colour of a point does not depend on the cell it is
computed in, a Gauss-Seidel colour pass with relax = 1 zeroes the residual on its colour, and the sweep from the
oracle's operator equals the sweep from the brute-force global matrix."""
import os

import numpy as np
import pytest

import oracle_lib as ol
import sor_model as sm
from bruteforce import RefinedMesh

MESHES = ["cube_6el.msh", "cube_24el.msh", "regular_octahedron_8el.msh", "three_tets_with_two_joint_faces.msh"]


def _setup(name, L):
    om = ol.Mesh(name)
    X = om.coords(0)[0]
    cells = om.cells()
    m, vcol = sm.vertex_colouring(len(X), cells)
    return om, X, cells, m, vcol


@pytest.mark.parametrize("name", MESHES)
def test_colouring_proper_on_micro_tets(name):
    L = 3
    om, X, cells, m, vcol = _setup(name, L)
    assert m >= 4
    bf = RefinedMesh(os.path.join(ol.MESHES, name), L)
    col = sm.point_colours(bf.xyz, X, cells, vcol, m, 1 << L)
    t = col[bf.tets]
    for a in range(4):
        for b in range(a + 1, 4):
            assert np.all(t[:, a] != t[:, b]), "two vertices of a micro-tet share a colour"


@pytest.mark.parametrize("name", MESHES)
def test_every_colour_used(name):
    L = 2
    om, X, cells, m, vcol = _setup(name, L)
    col = sm.point_colours(om.coords(L)[0], X, cells, vcol, m, 1 << L)
    assert set(np.unique(col)) == set(range(m))


def _oracle_parts(om, L, X, cells, m, vcol):
    op = ol.Operator(om, L, L)
    src, dst = ol.Function(om, L, L), ol.Function(om, L, L)

    def apply(x):
        src.set(L, x)
        op.apply(src, dst, L, ol.ALL, False)
        return dst.get(L)

    xyz, bnd = om.coords(L)
    col = sm.point_colours(xyz, X, cells, vcol, m, 1 << L)
    return apply, op.inv_diag(L), col, ~bnd, op


@pytest.mark.parametrize("name", ["cube_6el.msh", "cube_24el.msh"])
def test_gauss_seidel_pass_zeroes_its_colour(name):
    L = 3
    om, X, cells, m, vcol = _setup(name, L)
    apply, dinv, col, inner, _ = _oracle_parts(om, L, X, cells, m, vcol)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, len(col))
    b = rng.uniform(-1, 1, len(col))
    for k in range(m):
        # a single pass over colour k with relax = 1
        sel = inner & (col == k)
        r = b - apply(x)
        x[sel] += dinv[sel] * r[sel]
        r2 = b - apply(x)
        assert np.abs(r2[sel]).max() <= 1e-12 * np.abs(r).max()


@pytest.mark.parametrize("backwards", [False, True])
def test_sweep_oracle_equals_bruteforce(backwards):
    name, L = "cube_24el.msh", 3
    om, X, cells, m, vcol = _setup(name, L)
    apply, dinv, col, inner, _ = _oracle_parts(om, L, X, cells, m, vcol)
    xyz = om.coords(L)[0]
    bf = RefinedMesh(os.path.join(ol.MESHES, name), L)
    idx = bf.key_index(xyz)  # canonical DoF -> brute-force vertex
    A = bf.stiffness()[idx][:, idx]
    rng = np.random.default_rng(3)
    x0 = rng.uniform(-1, 1, len(col))
    b = rng.uniform(-1, 1, len(col))
    got = sm.sor_sweep(apply, dinv, x0, b, 1.3, col, m, inner, backwards)
    ref = sm.sor_sweep(lambda v: A @ v, 1.0 / A.diagonal(), x0, b, 1.3, col, m, inner, backwards)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
