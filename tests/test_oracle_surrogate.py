"""CPU pins of the oracle's surrogate operator (P1SurrogateOperator< p1_div_k_grad_affine_q3 >,
p1functionspace/P1SurrogateOperator.hpp) and of its polynomial coefficient.

* The reference's own theorem test (tests/hyteg/operators/P1SurrogateOperatorTest.cpp, 3D part): on the unit cube
  meshCuboid((0,0,0), (1,1,1), 1, 1, 1), for k a polynomial of degree p (the test's k1..k4: coefficient i + 1 on
  the i-th monomial) the surrogate of degree q = p interpolated on the apply level reproduces the variable
  operator: |A u - A_q u|_inf < 1e-12 for u = cos(2 pi x) cos(2 pi y) cos(2 pi z), flag All, cases (p, L) =
  (1, 3), (2, 3), (3, 3), (4, 3), (4, 4). The fitted stencils come from the oracle's point-wise stencil assembly
  and the reference apply from its element loops (pinned against a global assembly), so the test fails for a
  wrong stencil direction, sample point, basis order or fit. The inverse diagonals agree as well.
* The fit itself against numpy's least squares, with the monomial basis enumerated here from
  polynomial/MonomialBasis3D.hpp (graded; within a degree by decreasing power of x, then of y): Fourier k (not a
  polynomial), degree 3, on the apply level and one level below it.
* Without an exact representation the surrogate error falls as the degree grows.
"""
import numpy as np
import pytest

import oracle_lib as ol


def monomials(deg):
    out = []
    for d in range(deg + 1):
        for a in range(d, -1, -1):
            for b in range(d - a, -1, -1):
                out.append((a, b, d - a - b))
    return out


def poly_k(p):
    return ("poly", np.arange(1, len(monomials(p)) + 1, dtype=np.float64))


def interior_points(N):
    return [(x, y, z) for z in range(1, N - 2) for y in range(1, N - 1 - z) for x in range(1, N - y - z)]


@pytest.fixture(scope="module")
def unit_cube():
    return ol.Mesh("cuboid", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1, 1, 1)


def cosine(mesh, L):
    x, _ = mesh.coords(L)
    return np.cos(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]) * np.cos(2 * np.pi * x[:, 2])


@pytest.mark.parametrize("p,L", [(1, 3), (2, 3), (3, 3), (4, 3), (4, 4)])
def test_surrogate_of_degree_p_reproduces_a_degree_p_coefficient(unit_cube, p, L):
    m = unit_cube
    A = ol.DivKGradOperator(m, L, L, poly_k(p))
    Aq = ol.DivKGradOperator(m, L, L, poly_k(p))
    Aq.interpolate_stencils(p, L)
    u, au, aqu = ol.Function(m, L, L), ol.Function(m, L, L), ol.Function(m, L, L)
    u.set(L, cosine(m, L))
    A.apply(u, au, L, ol.ALL)
    Aq.apply(u, aqu, L, ol.ALL)
    err = np.abs(au.get(L) - aqu.get(L)).max()
    assert err < 1e-12, err
    assert np.abs(A.inv_diag(L) - Aq.inv_diag(L)).max() <= 1e-12 * np.abs(A.inv_diag(L)).max()


def test_a_lower_degree_does_not_reproduce_it(unit_cube):
    # the same check fails when q < p: the theorem test has teeth
    m, L = unit_cube, 3
    A = ol.DivKGradOperator(m, L, L, poly_k(3))
    Aq = ol.DivKGradOperator(m, L, L, poly_k(3))
    Aq.interpolate_stencils(1, L)
    u, au, aqu = ol.Function(m, L, L), ol.Function(m, L, L), ol.Function(m, L, L)
    u.set(L, cosine(m, L))
    A.apply(u, au, L, ol.ALL)
    Aq.apply(u, aqu, L, ol.ALL)
    assert np.abs(au.get(L) - aqu.get(L)).max() > 1e-6


@pytest.mark.parametrize("L,Li", [(4, 4), (5, 4)])
def test_fit_equals_numpy_least_squares(L, Li):
    m = ol.Mesh("cube_24el.msh")
    modes = ol.random_field(5, modes=4, kmax=1, std=0.5)
    deg = 3
    op = ol.DivKGradOperator(m, L, L, ("fourier", 1.1, modes))
    op.interpolate_stencils(deg, Li)
    Ni, sh = 2 ** Li, L - Li
    pts = interior_points(Ni)
    basis = monomials(deg)
    X = np.array(pts, dtype=np.float64) / Ni
    V = np.stack([X[:, 0] ** a * X[:, 1] ** b * X[:, 2] ** c for a, b, c in basis], axis=1)
    for c in (0, 7, 23):
        st = np.array([op.variable_stencil(L, c, x << sh, y << sh, z << sh) for x, y, z in pts])
        ref = np.linalg.lstsq(V, st, rcond=None)[0].T
        got = op.surrogate_poly(L, c)
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max(), (c, np.abs(got - ref).max())


def test_surrogate_error_falls_with_the_degree():
    # computeSurrogateError3D's quantity, the Frobenius norm of the stencil error over the interior points, on the
    # sample points themselves: least squares over nested polynomial spaces cannot increase it
    m, L = ol.Mesh("cube_24el.msh"), 4
    modes = ol.random_field(6, modes=4, kmax=1, std=0.5)
    N = 2 ** L
    pts = interior_points(N)
    X = np.array(pts, dtype=np.float64) / N
    errs = []
    for q in (0, 1, 2, 4):
        Aq = ol.DivKGradOperator(m, L, L, ("fourier", 1.0, modes))
        Aq.interpolate_stencils(q, L)
        V = np.stack([X[:, 0] ** a * X[:, 1] ** b * X[:, 2] ** c for a, b, c in monomials(q)], axis=1)
        e = 0.0
        for c in range(0, 24, 5):
            st = np.array([Aq.variable_stencil(L, c, *pt) for pt in pts])
            e += np.sum((V @ Aq.surrogate_poly(L, c).T - st) ** 2)
        errs.append(np.sqrt(e))
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
