"""Gmsh physical tags as mesh boundary flags (MeshInfo::fromGmshFile without setMeshBoundaryFlagsOnBoundary,
mesh/MeshGenFromGmshFile.cpp:100-116, 160-200) in the oracle, pinned by counting.

With the tagged cube of tests/gmsh_tags.py (x = 0 side tag 1, x = 1 side tag 2, y = 0 side tag 3, the four cube edges
of x = 0 tag 1) every line or triangle element gives its own edge or face the tag and everything else, the
diagonals of a side to its centre node and every vertex included, keeps flag 0. At level L (N = 2^L) that leaves:
flag 1 at the interior points of the x = 0 side's four triangles and of its four edges, 4 (N-1)(N-2)/2 + 4 (N-1)
points; flag 2 and flag 3 at the interior points of four triangles each, 4 (N-1)(N-2)/2 points; and without tags the
geometric rule of the untagged reader flags every boundary point 1.
"""
import os

import numpy as np
import pytest

import oracle_lib as ol
from gmsh_tags import write_tagged_cube


@pytest.mark.parametrize("L", [1, 2, 3, 4])
def test_tag_flags_by_count(tmp_path, L):
    path = os.path.join(str(tmp_path), "cube_tagged.msh")
    assert write_tagged_cube(path) == 4
    N = 2 ** L
    tri = (N - 1) * (N - 2) // 2
    m = ol.Mesh(path, tags=True)
    xyz, _ = m.coords(L)
    flags = np.zeros(m.ndofs(L), dtype=np.uint8)
    ol.lib().or_dof_coords(m.h, L, np.zeros(3 * m.ndofs(L)), flags)
    assert int((flags == 1).sum()) == 4 * tri + 4 * (N - 1)
    assert int((flags == 2).sum()) == 4 * tri
    assert int((flags == 3).sum()) == 4 * tri
    # the flagged points lie on their sides
    assert np.all(xyz[flags == 1][:, 0] == 0.0)
    assert np.all(xyz[flags == 2][:, 0] == 1.0)
    assert np.all(xyz[flags == 3][:, 1] == 0.0)
    # every macro-vertex keeps flag 0
    assert np.all(flags[:14] == 0)
    # the same file read with the geometric rule flags every boundary point
    g = ol.Mesh(path)
    gxyz, gb = g.coords(L)
    onb = np.any((gxyz == 0.0) | (gxyz == 1.0), axis=1)
    assert np.array_equal(gb, onb)


def test_tags_change_the_operator(tmp_path):
    """A Neumann-flagged side is not touched by an apply restricted to Inner and Dirichlet DoFs."""
    path = os.path.join(str(tmp_path), "cube_tagged.msh")
    write_tagged_cube(path)
    L = 3
    m = ol.Mesh(path, tags=True)
    n = m.ndofs(L)
    flags = np.zeros(n, dtype=np.uint8)
    ol.lib().or_dof_coords(m.h, L, np.zeros(3 * n), flags)
    op = ol.Operator(m, L, L)
    s, d = ol.Function(m, L, L), ol.Function(m, L, L)
    s.set(L, np.random.default_rng(3).uniform(-1, 1, n))
    d.set(L, np.full(n, 7.0))
    op.apply(s, d, L, ol.INNER | ol.DIRICHLET, False)
    out = d.get(L)
    assert np.all(out[flags == 2] == 7.0) and np.all(out[flags == 3] == 7.0)
    assert np.all(out[(flags == 0) | (flags == 1)] != 7.0)
