"""Gmsh physical tags as mesh boundary flags on the device (hyg_mesh_create_from_msh_tags) against the oracle.

On the tagged cube of tests/gmsh_tags.py (Dirichlet, Neumann and free-slip sides, untagged sides and every vertex
Inner): the Dirichlet mask equals the oracle's flag-1 points, the apply restricted to each DoF type matches the oracle
(1e-13 relative l-inf) and Jacobi V-cycles (Inner and Neumann DoFs solved) give the oracle's residual history
(1e-10 relative per cycle). The cycles start at level 2: below it the tagged cube has no Dirichlet point (vertices
and the side diagonals keep flag 0), so the coarse problem would be singular. Even so the cycles do not converge
on this mesh (the untagged sides, the vertices and the side diagonals are unknowns with one-sided rows; the oracle
grows the residual the same way), so the test checks that both follow the same iteration.
"""
import os

import numpy as np
import pytest

import oracle_lib as ol
from gmsh_tags import write_tagged_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tagged(hyg, tmp_path_factory):
    hyg.init()
    path = os.path.join(str(tmp_path_factory.mktemp("tags")), "cube_tagged.msh")
    write_tagged_cube(path)
    return hyg.Mesh(path, tags=True), ol.Mesh(path, tags=True)


def test_masks_and_apply_by_dof_type(hyg, tagged):
    gm, om = tagged
    for L in (2, 4):
        n = gm.ndofs(L)
        flags = np.zeros(n, dtype=np.uint8)
        ol.lib().or_dof_coords(om.h, L, np.zeros(3 * n), flags)
        assert np.array_equal(gm.boundary(L), flags == 1)
        src = np.random.default_rng(L).uniform(-1, 1, n)
        op, oop = hyg.P1ConstantOperator(gm, L, L), ol.Operator(om, L, L)
        for flag in (hyg.INNER, hyg.DIRICHLET, hyg.NEUMANN, hyg.FREESLIP, hyg.ALL):
            print(f"tagged cube L{L} apply flag {flag}", flush=True)
            s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
            s.upload(L, src)
            d.upload(L, np.full(n, 3.0))
            op.apply(s, d, L, flag)
            os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
            os_.set(L, src)
            od.set(L, np.full(n, 3.0))
            oop.apply(os_, od, L, flag, False)
            ref = od.get(L)
            assert np.abs(d.download(L) - ref).max() <= 1e-13 * np.abs(ref).max(), (L, flag)


def test_vcycle_history_on_the_tagged_cube(hyg, tagged):
    gm, om = tagged
    L = 4
    n = gm.ndofs(L)
    b0 = np.random.default_rng(9).uniform(-1, 1, n)
    b0[gm.boundary(L)] = 0.0
    kw = dict(pre=2, post=2, smoother="jacobi")
    minL = 2
    op = hyg.P1ConstantOperator(gm, minL, L)
    x, b = hyg.Function(gm, minL, L), hyg.Function(gm, minL, L)
    b.upload(L, b0)
    got = hyg.GMG(op, minL, L, **kw).solve(x, b, L, 4)
    oop = ol.Operator(om, minL, L)
    ox, ob = ol.Function(om, minL, L), ol.Function(om, minL, L)
    ob.set(L, b0)
    ref = ol.gmg(oop, ox, ob, minL, L, 4, **kw)
    assert np.all(np.abs(got - ref) / ref <= 1e-10), (got, ref)
