"""CPU checks that pin the P2 part of the oracle (oracle/oracle.cpp, test infrastructure).

A P2 function of level L is held on the vertex lattice of level L+1 (vertex DoFs at even lattice points,
edge DoFs at the midpoints of the level-L micro-edges). Every P2 routine of the oracle is compared with
something it shares no code with:
  * the P2 element matrix with exact integration of barycentric monomials (tests/bruteforce.py), and the
    global P2 stiffness matrix assembled from it element by element;
  * the quadratic interpolation matrix between levels, built from the P2 basis at the 35 lattice points of
    each coarse micro-tetrahedron;
  * the reference's own known-answer tests: tests/hyteg/P2/P2QuadraticProlongation3DTest.cpp (number of fine
    DoFs reached from one vertex / one Z edge DoF, exactness for quadratics) and
    tests/hyteg/P2/P2QuadraticRestriction3DTest.cpp (restriction of the constant 1: -7.5 at cell-interior
    vertex DoFs, 1 + 3/2 + n + n/4 + n/4 at cell-interior edge DoFs with n neighbour elements);
  * closed forms: A 1 = A x = 0 and A x^2 = -2 M 1 (M 1 = integrals of the basis functions, -vol/20 per
    micro-tetrahedron at a vertex, vol/5 at an edge) on inner DoFs.
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla
from scipy.spatial import cKDTree

import bruteforce as bf
import oracle_lib as ol

MESHDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "meshes")
MESHES = ["cube_24el.msh", "cube_kuhn_6el.msh", "tet_1el.msh", "pyramid_tilted_4el.msh",
          "three_tets_with_two_joint_faces.msh"]
# parity class of an edge orientation on the lattice, and its number of neighbour micro-cells
ORIENT = {"X": (1, 0, 0), "Y": (0, 1, 0), "Z": (0, 0, 1), "XY": (1, 1, 0), "XZ": (1, 0, 1), "YZ": (0, 1, 1),
          "XYZ": (1, 1, 1)}
NEIGHBOURS = {"X": 6, "Y": 4, "Z": 6, "XY": 6, "XZ": 4, "YZ": 6, "XYZ": 4}


def rnd(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


_models = {}


def model(name, L):
    k = (name, L)
    if k not in _models:
        _models[k] = bf.P2Model(os.path.join(MESHDIR, name), L)
    return _models[k]


def dof_map(om, bm, lat):
    """Index of each oracle DoF (lattice level lat) in the brute-force numbering."""
    xyz, bnd = om.coords(lat)
    d, idx = cKDTree(bm.lat.xyz).query(xyz)
    assert d.max() < 1e-9
    assert len(np.unique(idx)) == len(idx) == bm.n
    assert np.array_equal(bnd, bm.boundary[idx])
    return idx, bnd


def test_element_matrix_matches_exact_integration():
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (20, 4, 3))
    K = bf.p2_element_exact(X)
    for i in range(len(X)):
        A = ol.p2_element(X[i])
        assert rel_inf(A, K[i]) < 1e-13
        assert np.abs(A - A.T).max() < 1e-13 * np.abs(A).max()
        assert np.abs(A.sum(axis=1)).max() < 1e-13 * np.abs(A).max()


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("L", [1, 2])
def test_apply_matches_global_assembly(name, L):
    om = ol.Mesh(name)
    lat = L + 1
    bm = model(name, L)
    idx, bnd = dof_map(om, bm, lat)
    A = bm.stiffness()
    op = ol.P2Operator(om, L, L)
    n = om.ndofs(lat)
    src, old = rnd(n, 1), rnd(n, 2)
    s, d = ol.Function(om, lat, lat), ol.Function(om, lat, lat)
    s.set(lat, src)
    ug = np.zeros(bm.n)
    ug[idx] = src
    y = (A @ ug)[idx]
    for flag, mask in ((ol.INNER, ~bnd), (ol.ALL, np.ones(n, bool))):
        for add in (False, True):
            d.set(lat, old)
            op.apply(s, d, lat, flag, add)
            want = np.where(mask, y + (old if add else 0.0), old)
            assert rel_inf(d.get(lat), want) < 1e-13, (flag, add)


@pytest.mark.parametrize("name", ["cube_24el.msh", "cube_kuhn_6el.msh"])
def test_constants_linears_and_quadratics(name):
    """A 1 = A x = 0 on inner DoFs; A x^2 = -2 M 1 with M 1 the integrals of the P2 basis functions."""
    om = ol.Mesh(name)
    L = 2
    lat = L + 1
    bm = model(name, L)
    idx, bnd = dof_map(om, bm, lat)
    xyz, _ = om.coords(lat)
    op = ol.P2Operator(om, L, L)
    s, d = ol.Function(om, lat, lat), ol.Function(om, lat, lat)
    for u in (np.ones(len(xyz)), 3 * xyz[:, 0] - 2 * xyz[:, 1] + 0.5 * xyz[:, 2] + 1):
        s.set(lat, u)
        op.apply(s, d, lat, ol.INNER)
        assert np.abs(d.get(lat)[~bnd]).max() < 1e-12
    # integrals of the basis functions per micro-tetrahedron: -vol/20 (vertices), vol/5 (edges)
    J = bm.X[:, 1:, :] - bm.X[:, :1, :]
    vol = np.abs(np.linalg.det(J)) / 6.0
    m1 = np.zeros(bm.n)
    for k in range(10):
        np.add.at(m1, bm.nodes[:, k], vol * (-1.0 / 20.0 if k < 4 else 1.0 / 5.0))
    s.set(lat, xyz[:, 0] ** 2)
    op.apply(s, d, lat, ol.INNER)
    want = -2.0 * m1[idx]
    assert rel_inf(d.get(lat)[~bnd], want[~bnd]) < 1e-12


@pytest.mark.parametrize("name", MESHES)
def test_inverse_diagonal(name):
    om = ol.Mesh(name)
    L = 2
    bm = model(name, L)
    idx, bnd = dof_map(om, bm, L + 1)
    op = ol.P2Operator(om, L, L)
    dg = bm.stiffness().diagonal()[idx]
    assert rel_inf(op.inv_diag(L + 1), 1.0 / dg) < 1e-13


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("Lc", [0, 1, 2])
def test_transfers_match_interpolation_matrix(name, Lc):
    """Prolongation = quadratic interpolation matrix P (all DoFs), restriction = P^T (flag All)."""
    om = ol.Mesh(name)
    latc, latf = Lc + 1, Lc + 2
    cm, fm = model(name, Lc), model(name, Lc + 1)
    P = bf.prolongation_p2(cm, fm)
    ic, _ = dof_map(om, cm, latc)
    jf, bf_ = dof_map(om, fm, latf)
    f = ol.Function(om, latc, latf)
    u, old = rnd(len(ic), 3), rnd(len(jf), 4)
    ug = np.zeros(cm.n)
    ug[ic] = u
    pu = (P @ ug)[jf]
    for add in (False, True):
        f.set(latc, u)
        f.set(latf, old)
        ol.p2_prolongate(f, latc, ol.ALL, add)
        assert rel_inf(f.get(latf), pu + (old if add else 0.0)) < 1e-14
    # Inner | Neumann leaves Dirichlet DoFs untouched
    f.set(latc, u)
    f.set(latf, old)
    ol.p2_prolongate(f, latc, ol.INNER | ol.NEUMANN, False)
    assert rel_inf(f.get(latf), np.where(bf_, old, pu)) < 1e-14
    v = rnd(len(jf), 5)
    vg = np.zeros(fm.n)
    vg[jf] = v
    f.set(latf, v)
    ol.p2_restrict(f, latf, ol.ALL)
    assert rel_inf(f.get(latc), (P.T @ vg)[ic]) < 1e-13


@pytest.mark.parametrize("name", ["tet_1el.msh", "cube_24el.msh", "pyramid_tilted_4el.msh"])
def test_prolongation_is_exact_for_quadratics(name):
    """P2QuadraticProlongation3DTest.cpp testGridTransfer3D: 2x^2 + 3x + 13 + 4y + 5y^2 + z^2 + 6 to 1e-15."""
    om = ol.Mesh(name)
    q = lambda p: 2 * p[:, 0] ** 2 + 3 * p[:, 0] + 13 + 4 * p[:, 1] + 5 * p[:, 1] ** 2 + p[:, 2] ** 2 + 6  # noqa: E731
    for Lc in (0, 1, 2):
        xc, _ = om.coords(Lc + 1)
        xf, _ = om.coords(Lc + 2)
        f = ol.Function(om, Lc + 1, Lc + 2)
        f.set(Lc + 1, q(xc))
        ol.p2_prolongate(f, Lc + 1, ol.ALL, False)
        assert np.abs(f.get(Lc + 2) - q(xf)).max() < 1e-13 * np.abs(q(xf)).max()


def _cell_lattice(om, lat):
    """Canonical indices and cell-local lattice coordinates of the cell-interior DoFs of a 1-cell mesh."""
    xyz, bnd = om.coords(lat)
    nodes, tets = bf.read_gmsh(om.path)
    V = np.array([nodes[v] for v in tets[0]])
    N = 1 << lat
    loc = np.linalg.solve((V[1:] - V[0]).T, (xyz - V[0]).T).T * N
    ijk = np.rint(loc).astype(np.int64)
    assert np.abs(loc - ijk).max() < 1e-9
    inner = (ijk.min(axis=1) > 0) & (ijk.sum(axis=1) < N)
    return ijk, inner


def test_prolongation_reach_of_one_vertex_and_one_edge_dof():
    """P2QuadraticProlongation3DTest.cpp: on tet_1el, level 2 -> 3, vertex DoF (1,1,1) reaches
    1 + 2*14 + 3*24 + 24 fine DoFs; the Z edge DoF at (1,1,0) reaches 1 + 2 + 6*2 + 6 + 6."""
    om = ol.Mesh("tet_1el.msh")
    latc, latf = 3, 4
    ijk_c, _ = _cell_lattice(om, latc)
    ijk_f, inner_f = _cell_lattice(om, latf)
    for point, expect in (((2, 2, 2), 1 + 2 * 14 + 3 * 24 + 24), ((2, 2, 1), 1 + 2 + 6 * 2 + 6 * 1 + 6 * 1)):
        u = np.zeros(len(ijk_c))
        u[np.flatnonzero((ijk_c == point).all(axis=1))] = 1.0
        f = ol.Function(om, latc, latf)
        f.set(latc, u)
        ol.p2_prolongate(f, latc, ol.INNER | ol.NEUMANN, False)
        v = f.get(latf)
        assert int(np.sum(np.abs(v[inner_f]) > 1e-8)) == expect, point


@pytest.mark.parametrize("lower", [3, 4])
def test_restriction_of_one(lower):
    """P2QuadraticRestriction3DTest.cpp testWeightsInCell: restrict 1 from level lower+1 to lower on tet_1el."""
    om = ol.Mesh("tet_1el.msh")
    latc, latf = lower + 1, lower + 2
    ijk, inner = _cell_lattice(om, latc)
    f = ol.Function(om, latc, latf)
    f.set(latf, np.ones(om.ndofs(latf)))
    ol.p2_restrict(f, latf, ol.ALL)
    r = f.get(latc)
    par = ijk % 2
    vert = inner & (par == 0).all(axis=1)
    assert vert.any() and np.abs(r[vert] - (1 + 14 * 3 / 8 - 14 / 8 - 3 * 24 / 8 - 24 / 8)).max() < 1e-12
    for name, cls in ORIENT.items():
        sel = inner & (par == cls).all(axis=1)
        n = NEIGHBOURS[name]
        assert sel.any() and np.abs(r[sel] - (1 + 2 * 3 / 4 + n * 2 / 2 + n / 4 + n / 4)).max() < 1e-12, name


@pytest.mark.parametrize("name,smoother", [("cube_kuhn_6el.msh", "jacobi"), ("cube_24el.msh", "jacobi"),
                                           ("cube_24el.msh", "cheb")])
def test_vcycle_history_matches_assembled_model(name, smoother):
    """Residual norms per V-cycle of the oracle's P2 multigrid (P2 levels 0..2) vs the same cycle on
    assembled matrices and the interpolation matrix (inner DoFs)."""
    minL, maxL, cycles = 0, 2, 5
    om = ol.Mesh(name)
    gm = bf.P2GMGModel(os.path.join(MESHDIR, name), minL, maxL)
    idx, bnd = dof_map(om, gm.mesh[maxL], maxL + 1)
    n = om.ndofs(maxL + 1)
    b0 = rnd(n, 77)
    b0[bnd] = 0.0
    if smoother == "jacobi":
        kw = dict(pre=2, post=2, smoother="jacobi")
        smooth = lambda L, x, b: gm.jacobi(L, x, b, 2.0 / 3.0)  # noqa: E731
    else:
        kw = dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * 2.2, 1.2 * 2.2))
        coeffs = ol.cheb_coefficients(2, 0.3 * 2.2, 1.2 * 2.2)
        smooth = lambda L, x, b: gm.chebyshev(L, x, b, coeffs)  # noqa: E731
    op = ol.P2Operator(om, minL, maxL)
    x, b = ol.Function(om, minL + 1, maxL + 1), ol.Function(om, minL + 1, maxL + 1)
    b.set(maxL + 1, b0)
    res = ol.gmg(op, x, b, minL + 1, maxL + 1, cycles, **kw)
    inner = gm.inner[maxL]
    pos = np.empty(gm.mesh[maxL].n, dtype=np.int64)
    pos[idx] = np.arange(n)
    bi = b0[pos[inner]]
    xi = np.zeros(len(inner))
    ref = [np.linalg.norm(bi)]
    for _ in range(cycles):
        xi = gm.vcycle(maxL, xi, bi, smooth, kw["pre"], kw["post"])
        ref.append(np.linalg.norm(bi - gm.A[maxL] @ xi))
    ref = np.asarray(ref)
    assert ref[-1] < 0.05 * ref[0]
    assert np.all(np.abs(res - ref) <= 1e-10 * ref + 1e-14 * ref[0]), (res, ref)
    assert rel_inf(x.get(maxL + 1)[pos[inner]], xi) <= 1e-10


def test_coarse_solve_is_exact():
    """The V-cycle's coarsest level (P2 level 0 = lattice 1) is solved by CG to the direct solution."""
    name = "cube_24el.msh"
    om = ol.Mesh(name)
    gm = bf.P2GMGModel(os.path.join(MESHDIR, name), 0, 0)
    idx, bnd = dof_map(om, gm.mesh[0], 1)
    n = om.ndofs(1)
    b0 = rnd(n, 9)
    b0[bnd] = 0.0
    op = ol.P2Operator(om, 0, 0)
    x, b = ol.Function(om, 1, 1), ol.Function(om, 1, 1)
    b.set(1, b0)
    ol.gmg(op, x, b, 1, 1, 1)
    inner = gm.inner[0]
    pos = np.empty(gm.mesh[0].n, dtype=np.int64)
    pos[idx] = np.arange(n)
    xs = spla.spsolve(gm.A[0].tocsc(), b0[pos[inner]])
    assert rel_inf(x.get(1)[pos[inner]], xs) < 1e-12
