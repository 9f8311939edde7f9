"""Oracle restatement of MeshInfo::meshSymmetricCuboid (mesh/MeshGenSymmetricCuboid.cpp:35-253), pinned by what the
construction fixes: 24 tetrahedra per sub-cube around its 8 corners and 6 face centres, so the vertex count is the
corner lattice plus one centre per sub-cube face; the closed surface splits each sub-cube face into 4 triangles;
the complex is a ball (Euler characteristic 1); the cells fill the box exactly; every macro-cell has positive volume
and the boundary DoFs at level L lie on the box surface."""
import numpy as np
import pytest

import oracle_lib as ol

CASES = [((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1, 1, 1),
         ((-1.0, 0.5, 2.0), (2.0, 1.5, 2.75), 3, 1, 2),
         ((0.0, 0.0, 0.0), (1.0, 2.0, 3.0), 2, 2, 2)]


@pytest.mark.parametrize("case", CASES)
def test_counts_euler_volume(case):
    lo, hi, nx, ny, nz = case
    m = ol.Mesh("symmetric_cuboid", lo, hi, nx, ny, nz)
    nv, ne, nf, nc, bv, be, bf = [int(v) for v in m.info()]
    corners = (nx + 1) * (ny + 1) * (nz + 1)
    centres = nx * ny * (nz + 1) + ny * nz * (nx + 1) + nx * nz * (ny + 1)
    assert nv == corners + centres
    assert nc == 24 * nx * ny * nz
    assert nv - ne + nf - nc == 1
    assert bf == 4 * 2 * (nx * ny + ny * nz + nx * nz)
    # the surface is a sphere: V - E + F = 2
    assert bv - be + bf == 2
    vol = np.prod(np.array(hi) - np.array(lo))
    assert abs(m.volume() - vol) <= 1e-13 * vol


@pytest.mark.parametrize("case", CASES)
def test_boundary_dofs_on_the_box_surface(case):
    lo, hi, nx, ny, nz = case
    m = ol.Mesh("symmetric_cuboid", lo, hi, nx, ny, nz)
    lo, hi = np.array(lo), np.array(hi)
    for L in (0, 2):
        xyz, bnd = m.coords(L)
        tol = 1e-12 * np.abs(hi - lo).max()
        on_surface = (np.abs(xyz - lo) <= tol).any(axis=1) | (np.abs(xyz - hi) <= tol).any(axis=1)
        assert np.array_equal(bnd, on_surface)
        assert (xyz >= lo - tol).all() and (xyz <= hi + tol).all()


def test_single_cube_vertex_numbering():
    """The first sub-cube numbers its 14 vertices in local order (:103-107): corners of the bottom face, its
    centre, the four side centres, then the top face."""
    m = ol.Mesh("symmetric_cuboid", (0.0, 0.0, 0.0), (2.0, 2.0, 2.0), 1, 1, 1)
    xyz, _ = m.coords(0)
    want = [(0, 0, 0), (2, 0, 0), (1, 1, 0), (0, 2, 0), (2, 2, 0), (1, 0, 1), (0, 1, 1), (2, 1, 1), (1, 2, 1),
            (0, 0, 2), (2, 0, 2), (1, 1, 2), (0, 2, 2), (2, 2, 2)]
    assert np.array_equal(xyz[:14], np.array(want, dtype=float))


# MeshInfo::meshCuboid (mesh/MeshGenCuboid.cpp:35-174): six tetrahedra per grid cube, the grid points as vertices
CASES6 = [((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1, 1, 1), ((-1.0, 0.5, 2.0), (2.0, 1.5, 2.75), 3, 2, 2)]


@pytest.mark.parametrize("case", CASES6)
def test_cuboid_counts_euler_volume(case):
    lo, hi, nx, ny, nz = case
    m = ol.Mesh("cuboid", lo, hi, nx, ny, nz)
    nv, ne, nf, nc, bv, be, bf = [int(v) for v in m.info()]
    assert nv == (nx + 1) * (ny + 1) * (nz + 1)
    assert nc == 6 * nx * ny * nz
    assert nv - ne + nf - nc == 1
    # every boundary square of the grid is split into two triangles; the surface is a sphere
    assert bf == 2 * 2 * (nx * ny + ny * nz + nx * nz)
    assert bv - be + bf == 2
    vol = np.prod(np.array(hi) - np.array(lo))
    assert abs(m.volume() - vol) <= 1e-13 * vol


@pytest.mark.parametrize("case", CASES6)
def test_cuboid_operator_kills_linears(case):
    # P1LaplaceOperator3DTest.cpp:40-150 on this mesh: A applied to a linear function vanishes on inner DoFs
    m = ol.Mesh("cuboid", *case)
    L = 3
    xyz, bnd = m.coords(L)
    u = 0.3 + 1.7 * xyz[:, 0] - 0.4 * xyz[:, 1] + 2.2 * xyz[:, 2]
    op = ol.Operator(m, L, L)
    s, d = ol.Function(m, L, L), ol.Function(m, L, L)
    s.set(L, u)
    op.apply(s, d, L, ol.INNER, False)
    r = d.get(L)
    assert np.abs(r[~bnd]).max() <= 1e-11 * np.abs(u).max()


# MeshInfo::meshCubedDomain (mesh/MeshGenCubedDomain.cpp:37-383): an L-shaped block of unit cubes and a column
LSHAPE = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)]
COLUMN = [(-1, 2, 0), (-1, 2, 1), (-1, 2, 2)]


@pytest.mark.parametrize("cubes", [LSHAPE, COLUMN])
@pytest.mark.parametrize("cube_type", [0, 1])
def test_cubed_domain_counts_euler_volume(cubes, cube_type):
    m = ol.Mesh("cubed_domain", cubes, cube_type)
    nv, ne, nf, nc, bv, be, bf = [int(v) for v in m.info()]
    corners = len({(x + a, y + b, z + c) for x, y, z in cubes for a in (0, 1) for b in (0, 1) for c in (0, 1)})
    if cube_type == 0:
        assert nv == corners
    else:
        faces = {(2 * x + a, 2 * y + b, 2 * z + c) for x, y, z in cubes
                 for a, b, c in ((1, 1, 0), (1, 1, 2), (1, 0, 1), (1, 2, 1), (0, 1, 1), (2, 1, 1))}
        assert nv == corners + len(faces)
    assert nc == (6 if cube_type == 0 else 24) * len(cubes)
    assert nv - ne + nf - nc == 1
    assert bv - be + bf == 2
    assert abs(m.volume() - len(cubes)) <= 1e-13 * len(cubes)


@pytest.mark.parametrize("cube_type", [0, 1])
def test_cubed_domain_operator_kills_linears(cube_type):
    m = ol.Mesh("cubed_domain", LSHAPE, cube_type)
    L = 3
    xyz, bnd = m.coords(L)
    u = -0.7 + 0.9 * xyz[:, 0] + 1.3 * xyz[:, 1] - 2.1 * xyz[:, 2]
    op = ol.Operator(m, L, L)
    s, d = ol.Function(m, L, L), ol.Function(m, L, L)
    s.set(L, u)
    op.apply(s, d, L, ol.INNER, False)
    assert np.abs(d.get(L)[~bnd]).max() <= 1e-11 * np.abs(u).max()
