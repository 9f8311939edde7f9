"""Pins the positions of the oracle's spherical shell vertices (MeshInfo::meshSphericalShell, SHELLMESH_CLASSIC,
mesh/MeshGenSphericalShell.cpp) independently of its arc-midpoint recursion.

For ntan = 2^k + 1 the diamond bisection places the unit-sphere points of a k-times subdivided icosahedron whose
edge midpoints are pushed to the sphere (a class-I geodesic grid). Here that grid is built from scratch: the
icosahedron from the golden-ratio coordinates, its faces from the edge graph, k rounds of splitting every triangle
into four through normalised edge midpoints. The checks, each failing for a single misplaced point that still lies
on the sphere:
  * the mesh's unit directions equal the geodesic grid's up to a rotation: the sorted multisets of pairwise dot
    products agree (they are rotation invariant);
  * the mesh volume equals (rmax^3 - rmin^3) times the volume of the unit geodesic polyhedron (the shell between
    two scaled copies of it), every macro-cell has nonzero volume and the cells do not overlap (the sum of their
    absolute volumes is that same number).
"""
import itertools

import numpy as np
import pytest

import oracle_lib as ol


def geodesic(k):
    p = (1.0 + np.sqrt(5.0)) / 2.0
    v = []
    for a, b in itertools.product((-1.0, 1.0), repeat=2):
        v += [(0, a, b * p), (a, b * p, 0), (b * p, 0, a)]
    v = np.array(v) / np.linalg.norm(v[0])
    d = np.linalg.norm(v[:, None] - v[None, :], axis=2)
    edge = np.min(d[d > 1e-9])
    adj = np.abs(d - edge) < 1e-9
    faces = [f for f in itertools.combinations(range(12), 3) if adj[f[0], f[1]] and adj[f[1], f[2]] and adj[f[0], f[2]]]
    assert len(faces) == 20
    pts = [tuple(x) for x in v]
    tris = [tuple(v[i] for i in f) for f in faces]
    for _ in range(k):
        nxt = []
        for a, b, c in tris:
            ab, bc, ca = ((a + b), (b + c), (c + a))
            ab, bc, ca = ab / np.linalg.norm(ab), bc / np.linalg.norm(bc), ca / np.linalg.norm(ca)
            nxt += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        tris = nxt
    allp = np.array([x for t in tris for x in t])
    _, first = np.unique(np.round(allp, 10), axis=0, return_index=True)  # rounding only keys the duplicates
    uniq = allp[np.sort(first)]
    vol = sum(abs(np.linalg.det(np.array(t))) / 6.0 for t in tris)
    return uniq, vol


def dots(u):
    g = u @ u.T
    return np.sort(g[np.triu_indices(len(u), 1)])


def tet_volumes(xyz, cells):
    a = xyz[cells[:, 1]] - xyz[cells[:, 0]]
    b = xyz[cells[:, 2]] - xyz[cells[:, 0]]
    c = xyz[cells[:, 3]] - xyz[cells[:, 0]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0


@pytest.mark.parametrize("ntan,nrad,rmin,rmax", [(3, 2, 0.55, 1.0), (5, 2, 0.55, 1.0), (9, 3, 1.0, 2.0)])
def test_shell_vertices_are_the_geodesic_grid(ntan, nrad, rmin, rmax):
    k = int(round(np.log2(ntan - 1)))
    ref, vunit = geodesic(k)
    m = ol.Mesh("shell", ntan, nrad, rmin, rmax)
    xyz, _ = m.coords(0)
    r = np.linalg.norm(xyz, axis=1)
    outer = xyz[np.abs(r - rmax) < 1e-12] / rmax
    assert len(outer) == len(ref) == 10 * (ntan - 1) ** 2 + 2
    # every layer holds the same directions
    for rad in np.linspace(rmin, rmax, nrad):
        layer = xyz[np.abs(r - rad) < 1e-12] / rad
        assert np.abs(np.sort(layer, axis=0) - np.sort(outer, axis=0)).max() < 1e-13
    assert np.abs(dots(outer) - dots(ref)).max() < 1e-12
    # volumes
    vols = tet_volumes(xyz, m.cells())
    assert np.all(np.abs(vols) > 1e-6 * np.abs(vols).max())
    want = (rmax ** 3 - rmin ** 3) * vunit
    assert abs(np.abs(vols).sum() - want) < 1e-12 * want
    assert abs(m.volume() - want) < 1e-12 * want


def test_a_misplaced_point_is_caught():
    """The fingerprint changes when one point moves along the sphere (the check is not vacuous)."""
    ref, _ = geodesic(2)
    moved = ref.copy()
    t = moved[7] + 1e-3 * np.cross(moved[7], [0.3, 0.5, 0.8])
    moved[7] = t / np.linalg.norm(t)
    assert np.abs(dots(moved) - dots(ref)).max() > 1e-6
