"""CPU pins of the oracle's SOR in the reference's order (SORSmoothable::smooth_sor, P1Operator.hpp:343-416, with
VertexDoFMacroFace.hpp:751-810 and generatedKernels/sor_3D_macrocell_P1.cpp for the visiting order), restated in
oracle.cpp on the per-primitive memory with a snapshot standing for the ghost layers.

* On one macro-cell with only its interior updated (flag Inner), a forward sweep with relax = 1 is lexicographic
  Gauss-Seidel: x_I = (D + L)^-1 (b_I - U x_I - A_IB x_B), a triangular solve of the global matrix (brute-force
  assembly) in the lexicographic order.
* On multi-cell meshes, both directions, flags Inner and All, the oracle equals the global-matrix model of
  sor_model.ordered_sor (canonical order plus the closure rule for which neighbours are read before or after their
  update), written independently of the oracle's primitive loops and ghost handling.
* The closure rule matters: a plain Gauss-Seidel sweep in the same order (every neighbour current) differs.
* The reference's P1GMG3DConvergenceTest bound holds for Gauss-Seidel V(3,3) cycles with this smoother.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve_triangular

import oracle_lib as ol
import sor_model as sm
from bruteforce import RefinedMesh


_CACHE = {}


def global_matrix(name, om, L):
    if (name, L) in _CACHE:
        return _CACHE[(name, L)]
    xyz = om.coords(L)[0]
    bf = RefinedMesh(os.path.join(ol.MESHES, name), L)
    idx = bf.key_index(xyz)
    _CACHE[(name, L)] = bf.stiffness()[idx][:, idx].tocsr()
    return _CACHE[(name, L)]


def oracle_sweep(om, L, x0, b, relax, flag, backwards):
    op = ol.Operator(om, L, L)
    x, r = ol.Function(om, L, L), ol.Function(om, L, L)
    x.set(L, x0)
    r.set(L, b)
    assert ol.lib().or_smooth_sor_ordered(op.h, x.h, r.h, relax, L, flag, int(backwards)) == 0
    return x.get(L)


@pytest.mark.parametrize("L", [3, 4])
def test_one_cell_forward_sweep_is_a_triangular_solve(L):
    name = "tet_tilted_1el.msh"
    om = ol.Mesh(name)
    A = global_matrix(name, om, L)
    _, bnd = om.coords(L)
    rng = np.random.default_rng(11)
    x0, b = rng.uniform(-1, 1, len(bnd)), rng.uniform(-1, 1, len(bnd))
    got = oracle_sweep(om, L, x0, b, 1.0, ol.INNER, False)
    I, B = np.flatnonzero(~bnd), np.flatnonzero(bnd)
    AII = A[I][:, I]
    lower = sp.tril(AII, format="csr")
    upper = sp.triu(AII, k=1, format="csr")
    rhs = b[I] - upper @ x0[I] - A[I][:, B] @ x0[B]
    ref = x0.copy()
    ref[I] = spsolve_triangular(lower, rhs, lower=True)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name,L", [("cube_24el.msh", 3), ("cube_kuhn_6el.msh", 2), ("regular_octahedron_8el.msh", 2),
                                    ("three_tets_with_two_joint_faces.msh", 3)])
@pytest.mark.parametrize("flag", [ol.INNER, ol.ALL])
@pytest.mark.parametrize("backwards", [False, True])
def test_oracle_equals_the_global_matrix_model(name, L, flag, backwards):
    om = ol.Mesh(name)
    A = global_matrix(name, om, L)
    _, bnd = om.coords(L)
    prims = sm.canonical_primitives(om, L)
    update = np.ones(len(bnd), dtype=bool) if flag == ol.ALL else ~bnd
    rng = np.random.default_rng(5 + L)
    x0, b = rng.uniform(-1, 1, len(bnd)), rng.uniform(-1, 1, len(bnd))
    got = oracle_sweep(om, L, x0, b, 1.3, flag, backwards)
    ref = sm.ordered_sor(A, x0, b, 1.3, prims, update, backwards)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_the_closure_rule_matters():
    name, L = "cube_24el.msh", 3
    om = ol.Mesh(name)
    A = global_matrix(name, om, L)
    _, bnd = om.coords(L)
    prims = sm.canonical_primitives(om, L)
    rng = np.random.default_rng(9)
    x0, b = rng.uniform(-1, 1, len(bnd)), rng.uniform(-1, 1, len(bnd))
    ruled = sm.ordered_sor(A, x0, b, 1.0, prims, ~bnd)
    plain = sm.ordered_sor(A, x0, b, 1.0, prims, ~bnd, closure_rule=False)
    assert np.abs(ruled - plain).max() > 1e-3


def test_reference_gmg_convergence_bound():
    # tests/hyteg/P1GMG3DConvergenceTest.cpp:36-125 with this smoother: regular_octahedron_8el.msh, levels 0..3,
    # Gauss-Seidel V(3,3), u random in the interior, sin(x) sinh(y) z on the boundary, f = 0: the squared residual
    # falls by a factor below 3.2e-2 in every cycle
    om, L = ol.Mesh("regular_octahedron_8el.msh"), 3
    xyz, bnd = om.coords(L)
    u0 = np.random.default_rng(0).uniform(0.0, 1.0, len(xyz))
    u0[bnd] = np.sin(xyz[bnd, 0]) * np.sinh(xyz[bnd, 1]) * xyz[bnd, 2]
    x, b = ol.Function(om, 0, L), ol.Function(om, 0, L)
    x.set(L, u0)
    res = ol.gmg(ol.Operator(om, 0, L), x, b, 0, L, 4, pre=3, post=3, smoother="sor_ordered", relax=1.0)
    assert np.all((res[1:] / res[:-1]) ** 2 < 3.2e-2), res
