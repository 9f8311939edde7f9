"""CPU checks that pin the oracle's spherical shell mesh generator and its P2 variable-coefficient operator.

  * Shell (MeshInfo::meshSphericalShell, SHELLMESH_CLASSIC): the primitive counts the reference's own
    tests/hyteg/meshes/MeshGenTest.cpp expects for (ntan, nrad) = (5, 3) (486 vertices, 4160 faces, 1920 cells),
    the counts of (5, 2) (324, 1602, 2240, 960; boundary 324, 960, 640), and the volume of the shell (2, {1, 2})
    that tests/hyteg/operators/MassOperatorTest.cpp:197-200 derives in closed form (two icosahedra); every
    macro-vertex lies on one of the radii.
  * P2 div-k-grad element (p2_div_k_grad_affine_q4, Xiao-Gimbutas 11-point degree-4 rule): for the polynomial
    coefficient k = 1 + x/2 + y^2/4 the integrand has degree 4, so the element matrix must equal an independent
    collapsed-coordinate Gauss rule; a constant coefficient gives k0 times the (pinned) P2 Laplacian; the per-element
    factor xi_e comes from SplitMix64, whose first output for seed 0 is the published 0xe220a8397b1dcdaf.
"""
import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

import oracle_lib as ol

E = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]


def test_shell_primitive_counts():
    assert list(ol.Mesh("shell", 5, 3, 1.0, 2.0).info()[[0, 2, 3]]) == [486, 4160, 1920]
    assert list(ol.Mesh("shell", 5, 2, 0.55, 1.0).info()) == [324, 1602, 2240, 960, 324, 960, 640]


def test_shell_volume_and_radii():
    m = ol.Mesh("shell", 2, 2, 1.0, 2.0)
    a = 8.0 / np.sqrt(10.0 + 2.0 * np.sqrt(5.0))
    vol = 5.0 / 12.0 * (3.0 + np.sqrt(5.0)) * (a ** 3 - (a / 2) ** 3)
    assert abs(m.volume() - vol) < 1e-13 * vol
    for args in ((5, 2, 0.55, 1.0), (3, 3, 1.0, 2.0)):
        m = ol.Mesh("shell", *args)
        xyz, bnd = m.coords(0)
        r = np.linalg.norm(xyz, axis=1)
        radii = np.linspace(args[2], args[3], args[1])
        assert np.abs(r[:, None] - radii[None, :]).min(axis=1).max() < 1e-14
        assert np.array_equal(bnd, (np.abs(r - args[2]) < 1e-12) | (np.abs(r - args[3]) < 1e-12))


def _tet_rule(n):
    g, w = leggauss(n)
    g, w = (g + 1) / 2, w / 2
    pts, wts = [], []
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            for c, wc in zip(g, w):
                pts.append((a, b * (1 - a), c * (1 - a) * (1 - b)))
                wts.append(wa * wb * wc * (1 - a) ** 2 * (1 - b))
    return np.array(pts), np.array(wts)


def test_p2_divkgrad_element_is_exact_for_a_quadratic_coefficient():
    pts, wts = _tet_rule(6)
    rng = np.random.default_rng(0)
    for _ in range(10):
        X = rng.uniform(-1, 1, (4, 3))
        J = (X[1:] - X[0]).T
        det = abs(np.linalg.det(J))
        Jinv = np.linalg.inv(J)
        G = np.vstack([-(Jinv.sum(0)), Jinv])
        K = np.zeros((10, 10))
        for p, w in zip(pts, wts):
            lam = np.array([1 - p.sum(), p[0], p[1], p[2]])
            x = X[0] + J @ p
            k = 1 + 0.5 * x[0] + 0.25 * x[1] ** 2
            d = np.array([(4 * lam[v] - 1) * G[v] for v in range(4)] + [4 * (lam[a] * G[b] + lam[b] * G[a]) for a, b in E])
            K += w * det * k * d @ d.T
        assert np.abs(ol.p2_divkgrad_element_poly(X) - K).max() < 1e-13 * np.abs(K).max()


def test_splitmix_known_answer_and_coefficient():
    z = 0xE220A8397B1DCDAF
    xi = 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0
    x = np.array([0.3, -0.2, 0.5])
    assert ol.lib().or_p2k_value(0, 1.0, 0.0, 1.0, 0, x, 0) == pytest.approx(1.0 + xi, rel=1e-15)
    r = np.linalg.norm(x)
    v = ol.lib().or_p2k_value(0, 2.0, 2.99, 0.0, 7, x, 5)
    assert v == pytest.approx(2.0 * np.exp(2.99 * (1 - r)), rel=1e-14)


@pytest.mark.parametrize("name", ["cube_24el.msh", "three_tets_with_two_joint_faces.msh"])
def test_constant_coefficient_is_a_scaled_laplacian(name):
    om = ol.Mesh(name)
    L = 2
    lat = L + 1
    n = om.ndofs(lat)
    src = np.random.default_rng(1).uniform(-1, 1, n)
    s, d1, d2 = (ol.Function(om, lat, lat) for _ in range(3))
    s.set(lat, src)
    ol.P2Operator(om, L, L).apply(s, d1, lat, ol.INNER)
    ol.P2DivKGradOperator(om, L, L, kind=0, k0=2.5).apply(s, d2, lat, ol.INNER)
    ref = 2.5 * d1.get(lat)
    assert np.abs(d2.get(lat) - ref).max() < 1e-13 * np.abs(ref).max()
