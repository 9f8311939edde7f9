"""GPU parity of the surrogate operator (hyg_op_interpolate_stencils: P1SurrogateOperator<
p1_div_k_grad_affine_q3 >::interpolateStencils) and of the polynomial coefficient (hyg_op_create_p1_divkgrad_poly),
through the C ABI.

* The reference's theorem test (tests/hyteg/operators/P1SurrogateOperatorTest.cpp, 3D part) on the device: unit
  cube meshCuboid((0,0,0), (1,1,1), 1, 1, 1), k of degree p (coefficient i + 1 on the i-th monomial), surrogate of
  degree q = p interpolated on the apply level: |A u - A_q u|_inf < 1e-12 for u = cos(2 pi x) cos(2 pi y)
  cos(2 pi z), flag All; the variable operator with the polynomial coefficient matches the oracle's (1e-13).
* Apply (Inner / All, replace / add), inverse diagonal and Jacobi sweep of the surrogate with a Fourier coefficient
  against the oracle's surrogate, degrees 2 and 4, interpolation on the apply level and one level below: 1e-11
  relative l-inf (the two least-squares fits round differently; the polynomials are evaluated by Horner here and
  by direct sums in the oracle).
* A Jacobi V-cycle with the surrogate on every level: residual history against the oracle's (1e-9 per cycle).
* FP32 apply within 1e-5 of FP64; argument errors.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
# a smooth coefficient (k within about e^+-0.5 over the cube), which low-degree polynomial stencils approximate to a
# few per cent; rougher fields give fits whose centre weights cross zero, where any rounding difference is amplified
MODES = ol.random_field(41, modes=3, kmax=1, std=0.2)


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


def ncoef(p):
    return (p + 1) * (p + 2) * (p + 3) // 6


@pytest.fixture(scope="module")
def cubes(hyg):
    hyg.init()
    return {"unit": (hyg.Mesh("cuboid", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1, 1, 1),
                     ol.Mesh("cuboid", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1, 1, 1)),
            "cube_24el.msh": (hyg.Mesh("cube_24el.msh"), ol.Mesh("cube_24el.msh")),
            "cube_kuhn_6el.msh": (hyg.Mesh("cube_kuhn_6el.msh"), ol.Mesh("cube_kuhn_6el.msh"))}


@pytest.mark.parametrize("p,L", [(1, 3), (2, 3), (3, 3), (4, 3), (4, 4)])
def test_surrogate_reproduces_a_polynomial_coefficient(hyg, cubes, p, L):
    gm, om = cubes["unit"]
    k = np.arange(1, ncoef(p) + 1, dtype=np.float64)
    A = hyg.P1DivKGradOperator(gm, L, L, poly=k, poly_degree=p)
    Aq = hyg.P1DivKGradOperator(gm, L, L, poly=k, poly_degree=p)
    print(f"surrogate theorem p = q = {p}, level {L}", flush=True)
    Aq.interpolate_stencils(p, L)
    x, _ = gm.coords(L)
    u0 = np.cos(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]) * np.cos(2 * np.pi * x[:, 2])
    u, au, aqu = hyg.Function(gm, L, L), hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    u.upload(L, u0)
    A.apply(u, au, L, hyg.ALL)
    Aq.apply(u, aqu, L, hyg.ALL)
    assert np.abs(au.download(L) - aqu.download(L)).max() < 1e-12
    # the polynomial coefficient itself against the oracle
    oA = ol.DivKGradOperator(om, L, L, ("poly", k))
    ou, oau = ol.Function(om, L, L), ol.Function(om, L, L)
    ou.set(L, u0)
    oA.apply(ou, oau, L, ol.ALL)
    assert rel_inf(au.download(L), oau.get(L)) <= 1e-13


@pytest.mark.parametrize("name", ["cube_24el.msh", "cube_kuhn_6el.msh"])
@pytest.mark.parametrize("deg,L,Li", [(2, 4, 4), (4, 5, 5), (4, 5, 4)])
def test_surrogate_matches_the_oracle(hyg, cubes, name, deg, L, Li):
    gm, om = cubes[name]
    op = hyg.P1DivKGradOperator(gm, L, L, 1.2, MODES)
    print(f"surrogate {name} degree {deg} level {L} interpolation level {Li}", flush=True)
    op.interpolate_stencils(deg, Li)
    oop = ol.DivKGradOperator(om, L, L, ("fourier", 1.2, MODES))
    oop.interpolate_stencils(deg, Li)
    n = gm.ndofs(L)
    rng = np.random.default_rng(7 + L)
    src, d0, rhs = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
    os_.set(L, src)
    for flag in (hyg.INNER, hyg.ALL):
        for update in (hyg.REPLACE, hyg.ADD):
            s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
            s.upload(L, src)
            d.upload(L, d0)
            op.apply(s, d, L, flag, update)
            od.set(L, d0)
            oop.apply(os_, od, L, flag, update == hyg.ADD)
            assert rel_inf(d.download(L), od.get(L)) <= 1e-11, (flag, update)
    dinv = hyg.Function(gm, L, L)
    op.inverse_diagonal(dinv, L)
    assert rel_inf(dinv.download(L), oop.inv_diag(L)) <= 1e-11
    # Jacobi sweep
    s, b, out = hyg.Function(gm, L, L), hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    s.upload(L, src)
    b.upload(L, rhs)
    out.upload(L, src)  # points outside the flag keep their values, as in the oracle's copy below
    op.smooth_jac(out, b, s, 0.6, L, hyg.INNER)
    ob, oo = ol.Function(om, L, L), ol.Function(om, L, L)
    ob.set(L, rhs)
    oo.set(L, src)
    oop.smooth_jac(oo, ob, os_, 0.6, L, ol.INNER)
    assert rel_inf(out.download(L), oo.get(L)) <= 1e-11
    # FP32
    s32, d32 = hyg.Function(gm, L, L, hyg.F32), hyg.Function(gm, L, L, hyg.F32)
    s32.upload(L, src)
    op.apply(s32, d32, L, hyg.ALL)
    s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    s.upload(L, src)
    op.apply(s, d, L, hyg.ALL)
    assert rel_inf(d32.download(L), d.download(L)) <= 1e-5


def test_surrogate_vcycle_matches_the_oracle(hyg, cubes):
    gm, om = cubes["cube_24el.msh"]
    L, cycles = 5, 4
    op = hyg.P1DivKGradOperator(gm, 0, L, 1.2, MODES)
    op.interpolate_stencils(3, 4)
    oop = ol.DivKGradOperator(om, 0, L, ("fourier", 1.2, MODES))
    oop.interpolate_stencils(3, 4)
    n = gm.ndofs(L)
    _, bnd = gm.coords(L)
    b0 = np.random.default_rng(77).uniform(-1, 1, n)
    b0[bnd] = 0.0
    g = hyg.GMG(op, 0, L, pre=2, post=2, smoother="jacobi")
    x, b = hyg.Function(gm, 0, L), hyg.Function(gm, 0, L)
    b.upload(L, b0)
    print("surrogate V-cycles", flush=True)
    hist = g.solve(x, b, L, cycles)
    ox, ob = ol.Function(om, 0, L), ol.Function(om, 0, L)
    ob.set(L, b0)
    ref = ol.gmg(oop, ox, ob, 0, L, cycles, pre=2, post=2)
    assert np.all(np.abs(hist - ref) / ref <= 1e-9), (hist, ref)
    assert hist[-1] < 0.2 * hist[0]


def test_surrogate_argument_errors(hyg, cubes):
    gm, _ = cubes["cube_24el.msh"]
    op = hyg.P1DivKGradOperator(gm, 2, 3, 1.0, MODES)
    for deg, li in ((3, 1), (13, 3), (-1, 3)):
        with pytest.raises(hyg.HygError) as e:
            op.interpolate_stencils(deg, li)
        assert e.value.code == hyg.ERR_ARG
    const = hyg.P1ConstantOperator(gm, 2, 3)
    with pytest.raises(hyg.HygError) as e:
        check = hyg.load().hyg_op_interpolate_stencils(const.h, 2, 3)
        hyg.check(check)
    assert e.value.code == hyg.ERR_ARG
