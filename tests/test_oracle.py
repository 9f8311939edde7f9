"""CPU checks that pin the oracle (oracle/oracle.cpp, test infrastructure) before it is trusted.

Every oracle function is compared with something it does not share code with:
  * the global P1 stiffness matrix assembled element by element by tests/bruteforce.py
    (closed-form K_T = vol(T) G G^T, micro-tets from the lattice, vertices matched by coordinates);
  * the linear interpolation matrix between levels built from coarse micro-edges;
  * the reference's own known-answer tests for this path (tolerances and meshes quoted from
    tests/hyteg/P1/P1LaplaceOperator3DTest.cpp and
    tests/hyteg/vertexdofspace/VertexDoFLinearProlongation3DTest.cpp);
  * closed forms (Chebyshev polynomial identity, DoF counts of the configurations, h-scaling and
    symmetry of the stencil) and scipy solvers (sparse direct solve, symmetric eigenvalues).
A plausible mistake in the oracle (a dropped stencil term, a wrong orientation of a face map,
a transposed transfer, a wrong recurrence coefficient) fails at least one of these.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.spatial import cKDTree

import bruteforce as bf
import oracle_lib as ol

MESHDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "meshes")
BF_MESHES = ["cube_24el.msh", "cube_kuhn_6el.msh", "tet_1el.msh", "cube_6el.msh", "pyramid_tilted_4el.msh",
             "regular_octahedron_8el.msh", "three_tets_with_two_joint_faces.msh", "tet_tilted_1el.msh"]
KEY = 2 ** 20  # coordinate key resolution of the brute-force vertex identification


def rnd(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


_bf_cache = {}


def refined(name, L):
    k = (name, L)
    if k not in _bf_cache:
        _bf_cache[k] = bf.RefinedMesh(os.path.join(MESHDIR, name), L, scale_key=KEY)
    return _bf_cache[k]


def dof_map(om, bm, L):
    """Index of each oracle DoF in the brute-force numbering (one-to-one, by coordinates)."""
    xyz, bnd = om.coords(L)
    d, idx = cKDTree(bm.xyz).query(xyz)
    assert d.max() < 1e-9
    assert len(np.unique(idx)) == len(idx) == bm.n
    assert np.array_equal(bnd, bm.boundary[idx])
    return idx, bnd


# ---------------------------------------------------------------------------------------------- sizes
def test_dof_counts_of_the_configurations():
    """Vertex DoFs V + E(N-1) + F(N-1)(N-2)/2 + C(N-1)(N-2)(N-3)/6 and inner counts of the workloads."""
    m24, m6 = ol.Mesh("cube_24el.msh"), ol.Mesh("cube_kuhn_6el.msh")
    assert m24.ndofs(8) == 67_502_849
    assert m24.ndofs(7) == 8_487_297
    assert m6.ndofs(5) == 35_937
    _, b6 = m6.coords(5)
    assert int((~b6).sum()) == 29_791
    _, b24 = m24.coords(7)
    assert int((~b24).sum()) == 8_290_687
    # the Kuhn cube is a structured grid: (2^L + 1)^3 points, (2^L - 1)^3 inner
    for L in (2, 3, 4):
        _, b = m6.coords(L)
        assert len(b) == (2 ** L + 1) ** 3 and int((~b).sum()) == (2 ** L - 1) ** 3


# ---------------------------------------------------------------------------------------------- apply
@pytest.mark.parametrize("name", BF_MESHES)
@pytest.mark.parametrize("L", [1, 2, 3])
def test_apply_matches_assembled_matrix(name, L):
    om = ol.Mesh(name)
    bm = refined(name, L)
    idx, bnd = dof_map(om, bm, L)
    K = bm.stiffness()
    n = om.ndofs(L)
    x, d0 = rnd(n, 5 + L), rnd(n, 6 + L)
    xb = np.zeros(n)
    xb[idx] = x
    yfull = (K @ xb)[idx]
    op = ol.Operator(om, L, L)
    src, dst = ol.Function(om, L, L), ol.Function(om, L, L)
    src.set(L, x)
    for flag, add in ((ol.ALL, False), (ol.INNER, False), (ol.INNER, True)):
        dst.set(L, d0)
        op.apply(src, dst, L, flag, add)
        got = dst.get(L)
        ref = yfull if flag == ol.ALL else np.where(bnd, d0, (d0 if add else 0.0) + yfull)
        assert rel_inf(got, ref) <= 1e-13, (name, L, flag, add)


@pytest.mark.parametrize("name,L", [("tet_1el.msh", 2), ("tet_1el.msh", 3), ("pyramid_2el.msh", 2), ("pyramid_2el.msh", 3),
                                    ("pyramid_4el.msh", 3), ("pyramid_tilted_4el.msh", 3),
                                    ("regular_octahedron_8el.msh", 3)])
def test_laplace_annihilates_constants_and_linears(name, L):
    """P1LaplaceOperator3DTest.cpp:40-150: discrete L2 error of A u on Inner DoFs < 2.8e-13."""
    om = ol.Mesh(name)
    xyz, bnd = om.coords(L)
    op = ol.Operator(om, L, L)
    u, r = ol.Function(om, L, L), ol.Function(om, L, L)
    npts = int((~bnd).sum())
    for f in (np.zeros(len(xyz)), np.ones(len(xyz)), 42.0 * xyz[:, 0], 42.0 * xyz[:, 0] + xyz[:, 1] + 1337.0 * xyz[:, 2]):
        u.set(L, f)
        r.set(L, np.zeros(len(xyz)))
        op.apply(u, r, L, ol.INNER)
        err = np.sqrt(np.sum(r.get(L)[~bnd] ** 2) / npts)
        assert err < 2.8e-13, (name, L, err)


@pytest.mark.parametrize("name", ["cube_24el.msh", "pyramid_tilted_4el.msh"])
def test_stencil_scaling_and_symmetry(name):
    """3D P1 Laplace stencils scale with h (level L+1 = level L / 2), sum to zero and satisfy w(d) = w(-d)."""
    om = ol.Mesh(name)
    op = ol.Operator(om, 2, 5)
    opposite = [0, 2, 1, 4, 3, 6, 5, 11, 13, 12, 14, 7, 9, 8, 10]  # C W E N S NW SE TC TW TS TSE BC BN BE BNW
    for c in range(int(om.info()[3])):
        for L in (2, 3, 4):
            s, s1 = op.cell_stencil(L, c), op.cell_stencil(L + 1, c)
            assert rel_inf(s1, 0.5 * s) <= 1e-14
            assert abs(s.sum()) <= 1e-14 * np.abs(s).max()
            assert rel_inf(s[opposite], s) <= 1e-14


def test_inverse_diagonal_matches_matrix_diagonal():
    name, L = "cube_24el.msh", 3
    om = ol.Mesh(name)
    bm = refined(name, L)
    idx, _ = dof_map(om, bm, L)
    op = ol.Operator(om, L, L)
    assert rel_inf(op.inv_diag(L), 1.0 / bm.stiffness().diagonal()[idx]) <= 1e-14


# ---------------------------------------------------------------------------------------------- transfer
@pytest.mark.parametrize("name", BF_MESHES)
@pytest.mark.parametrize("Lf", [1, 2, 3])
def test_transfer_matches_interpolation_matrix(name, Lf):
    """Prolongation = linear interpolation P; restriction = P^T (flag All), restricted to the flagged DoFs."""
    Lc = Lf - 1
    om = ol.Mesh(name)
    cm, fm = refined(name, Lc), refined(name, Lf)
    ic, bc = dof_map(om, cm, Lc)
    jf, bfl = dof_map(om, fm, Lf)
    P = bf.prolongation(cm, fm)
    P = P[jf][:, ic]  # oracle numbering
    vc, vf = rnd(om.ndofs(Lc), 1), rnd(om.ndofs(Lf), 2)
    for flag in (ol.ALL, ol.INNER):
        for what in ("restrict", "prolong", "prolong_add"):
            f = ol.Function(om, Lc, Lf)
            f.set(Lc, vc)
            f.set(Lf, vf)
            if what == "restrict":
                ol.lib().or_restrict(f.h, Lf, flag)
                got, full, keep, old = f.get(Lc), P.T @ vf, bc, vc
            else:
                ol.lib().or_prolongate(f.h, Lc, flag, int(what == "prolong_add"))
                got, full, keep, old = f.get(Lf), P @ vc + (vf if what == "prolong_add" else 0.0), bfl, vf
            ref = full if flag == ol.ALL else np.where(keep, old, full)
            assert rel_inf(got, ref) <= 1e-14, (name, Lf, flag, what)


@pytest.mark.parametrize("name", ["tet_1el.msh", "pyramid_2el.msh", "pyramid_4el.msh", "pyramid_tilted_4el.msh",
                                  "regular_octahedron_8el.msh"])
@pytest.mark.parametrize("Lc", [0, 1, 3])
def test_prolongation_reproduces_linears(name, Lc):
    """VertexDoFLinearProlongation3DTest.cpp:45-136: sum of squared errors on Inner|Neumann < 1e-15."""
    om = ol.Mesh(name)
    xc, _ = om.coords(Lc)
    xf, bf_ = om.coords(Lc + 1)
    for fn in (lambda p: 0 * p[:, 0], lambda p: 1 + 0 * p[:, 0], lambda p: 42 + 0 * p[:, 0], lambda p: 42 * p[:, 0],
               lambda p: 42 * p[:, 0] + p[:, 1]):
        f = ol.Function(om, Lc, Lc + 1)
        f.set(Lc, fn(xc))
        ol.lib().or_prolongate(f.h, Lc, ol.INNER | ol.NEUMANN, 0)
        err = (f.get(Lc + 1) - fn(xf))[~bf_]
        assert np.sum(err ** 2) < 1e-15


def test_prolongate_and_add_equals_prolongate_plus_add():
    """VertexDoFLinearProlongation3DTest.cpp:138-195 (testProlongateAndAdd3D)."""
    for name in ("tet_1el.msh", "pyramid_4el.msh", "regular_octahedron_8el.msh"):
        om = ol.Mesh(name)
        for Lc in (0, 2):
            xc, _ = om.coords(Lc)
            xf, bnd = om.coords(Lc + 1)
            tp = np.sin(xc[:, 0]) + np.sinh(xc[:, 1]) + xc[:, 2] ** 2
            ta = np.cos(xf[:, 0]) + np.cosh(xf[:, 1]) + xf[:, 2] ** 3
            a, b = ol.Function(om, Lc, Lc + 1), ol.Function(om, Lc, Lc + 1)
            for f in (a, b):
                f.set(Lc, tp)
                f.set(Lc + 1, ta)
            ol.lib().or_prolongate(a.h, Lc, ol.INNER | ol.NEUMANN, 0)
            ol.lib().or_prolongate(b.h, Lc, ol.INNER | ol.NEUMANN, 1)
            pa = np.where(bnd, a.get(Lc + 1), a.get(Lc + 1) + ta)
            err = (pa - b.get(Lc + 1))[~bnd]
            assert np.sum(err ** 2) < 1e-15


# ---------------------------------------------------------------------------------------------- solvers
def test_chebyshev_coefficients_satisfy_the_polynomial_identity():
    """ChebyshevSmoother.hpp:128-190: 1 - lambda q(lambda) = T_k((theta - lambda)/delta) / T_k(theta/delta)."""
    lo, hi = 0.3 * 1.9, 1.2 * 1.9
    th, de = 0.5 * (hi + lo), 0.5 * (hi - lo)
    lam = np.linspace(0.0, 2.5, 41)
    for k in range(1, 6):
        c = ol.cheb_coefficients(k, lo, hi)
        q = sum(ci * lam ** i for i, ci in enumerate(c))
        Tk = np.polynomial.chebyshev.Chebyshev.basis(k)
        assert np.abs((1 - lam * q) - Tk((th - lam) / de) / Tk(th / de)).max() < 1e-12, k


def _bf_gmg(name, minL, maxL):
    return bf.GMGModel(os.path.join(MESHDIR, name), minL, maxL)


@pytest.mark.parametrize("name,smoother", [("cube_kuhn_6el.msh", "jacobi"), ("cube_24el.msh", "jacobi"),
                                           ("cube_24el.msh", "cheb")])
def test_vcycle_history_matches_assembled_model(name, smoother):
    """Per-cycle residual norms of the oracle V-cycle vs the same cycle on assembled matrices (inner DoFs)."""
    minL, maxL, cycles = 0, 3, 5
    om = ol.Mesh(name)
    model = _bf_gmg(name, minL, maxL)
    idx, bnd = dof_map(om, model.mesh[maxL], maxL)
    n = om.ndofs(maxL)
    b0 = rnd(n, 77)
    b0[bnd] = 0.0
    if smoother == "jacobi":
        kw = dict(pre=2, post=2, smoother="jacobi")
        smooth = lambda L, x, b: model.jacobi(L, x, b, 2.0 / 3.0)  # noqa: E731
    else:
        kw = dict(pre=3, post=3, smoother="cheb", cheb_order=3, cheb_bounds=(0.3 * 1.9, 1.2 * 1.9))
        coeffs = ol.cheb_coefficients(3, 0.3 * 1.9, 1.2 * 1.9)
        smooth = lambda L, x, b: model.chebyshev(L, x, b, coeffs)  # noqa: E731
    op = ol.Operator(om, minL, maxL)
    x, b = ol.Function(om, minL, maxL), ol.Function(om, minL, maxL)
    b.set(maxL, b0)
    res = ol.gmg(op, x, b, minL, maxL, cycles, **kw)
    # the same cycle on the assembled model, in the model's inner-DoF numbering
    inner = model.inner[maxL]
    pos = np.empty(model.mesh[maxL].n, dtype=np.int64)
    pos[idx] = np.arange(n)
    bi = b0[pos[inner]]
    xi = np.zeros(len(inner))
    ref = [np.linalg.norm(bi)]
    for _ in range(cycles):
        xi = model.vcycle(maxL, xi, bi, smooth, kw["pre"], kw["post"])
        ref.append(np.linalg.norm(bi - model.A[maxL] @ xi))
    # 1e-10 relative per cycle, with a floor of 1e-14 |b|: once the residual has dropped by
    # ~1e-8 its rounding error (different summation orders, ~eps |b|) is no longer below 1e-10 of it
    ref = np.asarray(ref)
    assert np.all(np.abs(res - ref) <= 1e-10 * ref + 1e-14 * ref[0]), (res, ref)
    xo = x.get(maxL)[pos[inner]]
    assert rel_inf(xo, xi) <= 1e-10


def test_coarse_cg_solves_the_system():
    """CGSolver (identity preconditioner) on a single level reproduces the sparse direct solution."""
    name, L = "cube_24el.msh", 2
    om = ol.Mesh(name)
    model = _bf_gmg(name, L, L)
    idx, bnd = dof_map(om, model.mesh[L], L)
    n = om.ndofs(L)
    b0 = rnd(n, 3)
    b0[bnd] = 0.0
    op = ol.Operator(om, L, L)
    x, b = ol.Function(om, L, L), ol.Function(om, L, L)
    b.set(L, b0)
    ol.gmg(op, x, b, L, L, 1, cg_tol=1e-14)
    pos = np.empty(n, dtype=np.int64)
    pos[idx] = np.arange(n)
    inner = model.inner[L]
    ref = spla.spsolve(model.A[L].tocsc(), b0[pos[inner]])
    assert rel_inf(x.get(L)[pos[inner]], ref) <= 1e-10


def test_power_iteration_estimates_the_spectral_radius():
    """estimateSpectralRadiusWithPowerIteration on D^-1 A: below and close to the largest eigenvalue."""
    name, L = "cube_24el.msh", 3
    om = ol.Mesh(name)
    bm = refined(name, L)
    idx, _ = dof_map(om, bm, L)
    K = bm.stiffness()
    d = K.diagonal()
    # D^-1 K is similar to the symmetric D^-1/2 K D^-1/2
    S = sp.diags(d ** -0.5) @ K @ sp.diags(d ** -0.5)
    rho = spla.eigsh(S, k=1, which="LA", return_eigenvectors=False)[0]
    op = ol.Operator(om, L, L)
    v = ol.Function(om, L, L)
    v.set(L, rnd(om.ndofs(L), 5))
    est = ol.lib().or_power_iteration(op.h, v.h, 40, L)
    assert est <= rho * (1 + 1e-12)
    assert est >= 0.9 * rho


def test_mixed_precision_refinement_reaches_the_fp64_solution():
    """Defect correction with FP32 V-cycles converges to the FP64 solution (cf. the 5e-9 check of
    tests/hyteg/mixedPrecision/mixedPrecisionIterativeRefinement.cpp:223-224)."""
    name, minL, maxL = "cube_kuhn_6el.msh", 0, 4
    om = ol.Mesh(name)
    n = om.ndofs(maxL)
    _, bnd = om.coords(maxL)
    b0 = rnd(n, 9)
    b0[bnd] = 0.0
    kw = dict(pre=2, post=2, smoother="jacobi")
    opd, ops = ol.Operator(om, minL, maxL, 8), ol.Operator(om, minL, maxL, 4)
    xd, bd = ol.Function(om, minL, maxL), ol.Function(om, minL, maxL)
    bd.set(maxL, b0)
    res = ol.gmg(opd, xd, bd, minL, maxL, 40, **kw)
    assert res[-1] / res[0] < 1e-12
    xm = ol.Function(om, minL, maxL)
    ip = np.array([minL, maxL, 2, 2, 0, 0, 2, 10000], dtype=np.int32)
    dp = np.array([2.0 / 3.0, 0.0, 0.0, 1e-16])
    rir = np.zeros(41)
    ol.lib().or_ir(opd.h, ops.h, xm.h, bd.h, ip, dp, 40, 1, rir)
    assert rir[-1] / rir[0] < 1e-12
    assert rel_inf(xm.get(maxL), xd.get(maxL)) < 5e-9
