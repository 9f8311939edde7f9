"""CPU checks that pin the oracle's variable-coefficient operator -div(k grad u)
(P1ElementwiseOperator<p1_div_k_grad_affine_q3>), test infrastructure for the GPU parity tests.

  * k constant: the elementwise operator equals k times the constant-coefficient stencil operator
    (tests/hyteg/operators/ElementwiseOperatorAdditiveApplyTest.cpp:113-125);
  * k variable: equals the global matrix assembled element by element with a numpy restatement of the
    5-point degree-3 rule (centroid weight -4/5, four points (1/2,1/6,1/6,1/6) with weight 9/20, times
    the element volume);
  * the reference's convergence test (tests/hyteg/convergence/ElementwiseDivKGradCGConvergenceTest.cpp:
    tet_1el levels 3/4/5, discrete L2 errors below 1.1e-3 / 3.5e-4 / 1.0e-4).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.spatial import cKDTree

import bruteforce as bf
import oracle_lib as ol

MESHDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "meshes")


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


def kfun_fourier(modes, k0=1.0):
    def k(p):
        g = np.zeros(p.shape[:-1])
        for kx, ky, kz, ph, a in modes:
            g += a * np.cos(2 * np.pi * (kx * p[..., 0] + ky * p[..., 1] + kz * p[..., 2]) + ph)
        return k0 * np.exp(g)
    return k


def quad_factor(k):
    """per-element factor for RefinedMesh.stiffness: (sum_q w_q k(x_q)) / volume"""
    lam = np.array([[0.25] * 4] + [[0.5 if i == j else 1.0 / 6 for i in range(4)] for j in range(4)])
    w = np.array([-0.8, 0.45, 0.45, 0.45, 0.45])

    def f(X):  # X: (ne, 4, 3)
        pts = np.einsum("qv,evc->eqc", lam, X)
        return (k(pts) * w[None, :]).sum(axis=1)
    return f


def dof_map(om, bm, L):
    xyz, bnd = om.coords(L)
    d, idx = cKDTree(bm.xyz).query(xyz)
    assert d.max() < 1e-9 and len(np.unique(idx)) == bm.n
    return idx, bnd


@pytest.mark.parametrize("name", ["cube_24el.msh", "tet_tilted_1el.msh", "pyramid_tilted_4el.msh"])
def test_constant_k_equals_scaled_stencil_operator(name):
    om = ol.Mesh(name)
    L = 3
    n = om.ndofs(L)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    ek = ol.DivKGradOperator(om, L, L, ("fourier", 2.5, np.zeros((0, 5))))
    ck = ol.Operator(om, L, L)
    s, a, b = ol.Function(om, L, L), ol.Function(om, L, L), ol.Function(om, L, L)
    s.set(L, x)
    for flag in (ol.INNER, ol.ALL):
        a.set(L, np.zeros(n))
        b.set(L, np.zeros(n))
        ek.apply(s, a, L, flag)
        ck.apply(s, b, L, flag)
        assert rel_inf(a.get(L), 2.5 * b.get(L)) <= 1e-13
    assert rel_inf(ek.inv_diag(L), ck.inv_diag(L) / 2.5) <= 1e-13


@pytest.mark.parametrize("name", ["cube_24el.msh", "cube_kuhn_6el.msh", "pyramid_tilted_4el.msh", "three_tets_with_two_joint_faces.msh"])
@pytest.mark.parametrize("L", [1, 2, 3])
def test_variable_k_matches_assembled_matrix(name, L):
    om = ol.Mesh(name)
    bm = bf.RefinedMesh(os.path.join(MESHDIR, name), L, scale_key=2 ** 20)
    idx, bnd = dof_map(om, bm, L)
    modes = ol.random_field(11, modes=8, kmax=3)
    K = bm.stiffness(quad_factor(kfun_fourier(modes, 1.3)))
    n = om.ndofs(L)
    x, d0 = np.random.default_rng(L).uniform(-1, 1, n), np.random.default_rng(9 + L).uniform(-1, 1, n)
    xb = np.zeros(n)
    xb[idx] = x
    ref = (K @ xb)[idx]
    op = ol.DivKGradOperator(om, L, L, ("fourier", 1.3, modes))
    s, d = ol.Function(om, L, L), ol.Function(om, L, L)
    s.set(L, x)
    for flag, add in ((ol.ALL, False), (ol.INNER, False), (ol.INNER, True)):
        d.set(L, d0)
        op.apply(s, d, L, flag, add)
        exp = ref if flag == ol.ALL else np.where(bnd, d0, (d0 if add else 0.0) + ref)
        assert rel_inf(d.get(L), exp) <= 1e-13, (name, L, flag, add)
    assert rel_inf(op.inv_diag(L), 1.0 / K.diagonal()[idx]) <= 1e-13


def mass_matrix(bm):
    X = bm.xyz[bm.tets]
    vol = np.abs(np.linalg.det(X[:, 1:, :] - X[:, :1, :])) / 6.0
    loc = (np.ones((4, 4)) + np.eye(4)) / 20.0
    M = vol[:, None, None] * loc[None]
    rows = np.repeat(bm.tets, 4, axis=1).reshape(-1)
    cols = np.tile(bm.tets, (1, 4)).reshape(-1)
    return sp.csr_matrix((M.reshape(-1), (rows, cols)), shape=(bm.n, bm.n))


@pytest.mark.parametrize("L,target", [(3, 1.1e-3), (4, 3.5e-4), (5, 1.0e-4)])
def test_reference_cg_convergence_on_tet(L, target):
    """ElementwiseDivKGradCGConvergenceTest.cpp:49-143 (P1, tet_1el): CG (tol 1e-12, 1000 its) on
    -div(k grad u) = f, k = tanh(2(x-1/2)) + 2, u = sin(2 pi x) sinh(pi y), f = M r."""
    name, alpha, phi = "tet_1el.msh", 2.0, 2.0
    om = ol.Mesh(name)
    bm = bf.RefinedMesh(os.path.join(MESHDIR, name), L, scale_key=2 ** 20)
    idx, bnd = dof_map(om, bm, L)
    xyz, _ = om.coords(L)
    exact = np.sin(phi * np.pi * xyz[:, 0]) * np.sinh(np.pi * xyz[:, 1])
    t0 = np.tanh(alpha * (xyz[:, 0] - 0.5))
    t1 = phi * alpha * (t0 * t0 - 1) * np.cos(phi * np.pi * xyz[:, 0])
    t2 = (phi * phi - 1) * np.pi * (t0 + 2) * np.sin(phi * np.pi * xyz[:, 0])
    r = np.pi * (t1 + t2) * np.sinh(np.pi * xyz[:, 1])
    rb = np.zeros(bm.n)
    rb[idx] = r
    f = (mass_matrix(bm) @ rb)[idx]
    op = ol.DivKGradOperator(om, L, L, ("tanh", alpha))
    u, b = ol.Function(om, L, L), ol.Function(om, L, L)
    u.set(L, np.where(bnd, exact, 0.0))
    b.set(L, f)
    ol.gmg(op, u, b, L, L, 1, cg_max_iter=1000, cg_tol=1e-12)
    err = u.get(L) - exact
    npoints = ol.Mesh(name).ndofs(L + 1)  # numberOfGlobalDoFs<P2FunctionTag>(level)
    e = np.sqrt(np.sum(err ** 2) / npoints)
    assert e < target, (L, e)
