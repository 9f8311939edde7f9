"""CPU pins of the oracle's spherical-shell blending (IcosahedralShellMap, geometry/IcosahedralShellMap.hpp) and of its
blended P1 Laplacian (P1ElementwiseBlendingLaplaceOperator = P1ElementwiseOperator< p1_diffusion_blending_q3 >),
against tests/shell_map.py, written separately from the definitions:

* the per-cell map parameters (ray pair, lone vertex, skew cells taking a neighbour's map) agree;
* the map's Jacobian (derived in closed form in both) matches central differences of the map;
* the map keeps the macro-vertices, sends the points of a cell's inner and outer faces onto the spheres, and is
  continuous: a face point mapped with either neighbour cell's map lands on the same position;
* the blended shell volume is (4 pi / 3)(r_max^3 - r_min^3) up to the quadrature error, which falls with the level
  (the affine macro-mesh misses it by far more);
* the oracle's operator equals a global assembly of the restated element matrices over the independently refined
  micro-mesh (tests/bruteforce.py), and its blended DoF coordinates are the map of the affine ones.
"""
import numpy as np
import pytest

import oracle_lib as ol
import shell_map as sm
from bruteforce import MICRO_TETS, RefinedMesh

SHELLS = [(3, 2, 0.55, 1.0), (5, 2, 0.55, 1.0), (3, 3, 0.5, 1.0)]


def shell(args):
    m = ol.Mesh("shell", *args)
    X = m.coords(0)[0]
    cells = m.cells()
    return m, X, cells


@pytest.mark.parametrize("args", SHELLS)
def test_map_parameters_agree(args):
    m, X, cells = shell(args)
    m.set_shell_map()
    ref = sm.classify(X, cells)
    for c in range(len(cells)):
        got = m.shell_map(c)
        for a, b in zip(got, ref[c]):
            assert np.allclose(a, b, rtol=0, atol=1e-15)


@pytest.mark.parametrize("args", SHELLS[:2])
def test_jacobian_and_map_properties(args):
    m, X, cells = shell(args)
    P = sm.classify(X, cells)
    rng = np.random.default_rng(1)
    rmin, rmax = args[2], args[3]
    faces = {}
    for c, cv in enumerate(cells):
        for f in sm.LOCAL_FACES:
            faces.setdefault(tuple(sorted(int(cv[i]) for i in f)), []).append(c)
    for c, cv in enumerate(cells):
        V = X[cv]
        assert np.abs(sm.F(P[c], V) - V).max() <= 1e-14
        x = rng.dirichlet(np.ones(4), 4) @ V
        D = sm.DF(P[c], x)
        h = 1e-6
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = (sm.F(P[c], x + e) - sm.F(P[c], x - e)) / (2 * h)
            assert np.abs(fd - D[:, :, k]).max() <= 1e-8
    for ids, cs in faces.items():
        V = X[list(ids)]
        pts = rng.dirichlet(np.ones(3), 5) @ V
        r = np.linalg.norm(V, axis=1)
        if np.allclose(r, rmin) or np.allclose(r, rmax):
            assert np.allclose(np.linalg.norm(sm.F(P[cs[0]], pts), axis=1), r[0], rtol=0, atol=1e-14)
        if len(cs) == 2:
            assert np.abs(sm.F(P[cs[0]], pts) - sm.F(P[cs[1]], pts)).max() <= 1e-14


def blended_volume(X, cells, P, L):
    N = 1 << L
    lat = np.array([(x, y, z) for z in range(N + 1) for y in range(N + 1 - z) for x in range(N + 1 - y - z)])
    vol = 0.0
    for c, cv in enumerate(cells):
        V = X[cv]
        for t in MICRO_TETS:
            verts = lat[:, None, :] + t[None, :, :]
            ok = np.all(verts.sum(axis=2) <= N, axis=1)
            for vv in verts[ok]:
                Xe = V[0] + (vv @ (V[1:] - V[0])) / N
                J = (Xe[1:] - Xe[0]).T
                for lam, w in zip(sm.QUAD_LAM, sm.QUAD_W):
                    vol += abs(np.linalg.det(J)) * w * abs(np.linalg.det(sm.DF(P[c], lam @ Xe)[0]))
    return vol


def test_blended_volume_converges():
    args = (3, 2, 0.55, 1.0)
    _, X, cells = shell(args)
    P = sm.classify(X, cells)
    exact = 4.0 / 3.0 * np.pi * (args[3] ** 3 - args[2] ** 3)
    e = [abs(blended_volume(X, cells, P, L) - exact) for L in (0, 1, 2)]
    affine = abs(ol.Mesh("shell", *args).volume() - exact)  # the affine macro-mesh misses the curvature
    assert e[0] < affine / 50 and e[1] < e[0] and e[2] < e[0] / 10, (e, affine)


@pytest.mark.parametrize("args,L", [((3, 2, 0.55, 1.0), 2), ((5, 2, 0.55, 1.0), 1), ((3, 3, 0.5, 1.0), 2)])
def test_operator_equals_the_global_assembly(args, L):
    m, X, cells = shell(args)
    xyz_affine = m.coords(L)[0]
    m.set_shell_map()
    P = sm.classify(X, cells)
    bf = RefinedMesh(({i: X[i] for i in range(len(X))}, [list(c) for c in cells]), L, scale_key="tree")
    per_cell = (1 << L) ** 3
    rows, cols, vals = [], [], []
    for e, tet in enumerate(bf.tets):
        A = sm.element_matrix(P[e // per_cell], bf.xyz[tet])
        rows.append(np.repeat(tet, 4))
        cols.append(np.tile(tet, 4))
        vals.append(A.reshape(-1))
    import scipy.sparse as sp
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(bf.n, bf.n))
    idx = bf.index_of(xyz_affine)
    K = K[idx][:, idx]
    u = np.random.default_rng(4).uniform(-1, 1, len(idx))
    op = ol.BlendingLaplaceOperator(m, L, L)
    s, d = ol.Function(m, L, L), ol.Function(m, L, L)
    s.set(L, u)
    op.apply(s, d, L, ol.ALL)
    ref = K @ u
    assert np.abs(d.get(L) - ref).max() <= 1e-12 * np.abs(ref).max()
    # blended coordinates: the map of the affine ones (vertices unchanged)
    xyz_blended = m.coords(L)[0]
    nv = len(X)
    assert np.abs(xyz_blended[:nv] - xyz_affine[:nv]).max() == 0.0
    r = np.linalg.norm(xyz_blended, axis=1)
    assert r.min() >= args[2] - 1e-14 and r.max() <= args[3] + 1e-14
