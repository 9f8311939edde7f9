"""GPU parity of the solver layer against the oracle, through the C ABI.

  * hyg_power_iteration (chebyshev::estimateRadius -> estimateSpectralRadiusWithPowerIteration,
    solvers/ChebyshevSmoother.hpp:255-265, numerictools/SpectrumEstimation.hpp:57-80): same start vector, 40
    iterations, radius within 1e-12 relative of the oracle's (whose estimate is pinned against scipy's eigsh).
  * hyg_ir_solve (tests/hyteg/mixedPrecision/mixedPrecisionIterativeRefinement.cpp:162-172, checked at :223-224):
    FP32 V-cycles inside FP64 defect correction, P1 constant and div-k-grad operators; the final iterate lies within
    5e-9 (relative l-inf) of the FP64 GMG solution of the oracle, and the number of outer steps to 1e-10 equals the
    oracle's or_ir count within one step (the FP32 inner cycles round in a different order on the device).
  * hyg_gmg_vcycle_graph on meshes whose coarsest level has inner DoFs, so the device-side coarse CG takes steps
    (13 inner DoFs at level 1 of the 24-tet cube; the 3 x 1 x 2 symmetric cuboid at level 0; P2 level 0): every
    replayed cycle equals the oracle's V-cycle (GeometricMultigridSolver.hpp:136-235) within 1e-10, with eager
    work between replays; SOR and symmetric SOR cycles captured in a graph equal the eager cycles.
  * the configuration-5 operator (P2 div-k-grad, eta(r) = exp(2.99 (1 - r)) (1 + 0.1 xi)) on the spherical shell
    (5, 2): apply, inverse diagonal within 1e-13 and Chebyshev V(3,3) residual history within 1e-10 per cycle.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
TOL64 = 1e-13


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


def rnd(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


FIELD = None


def field():
    global FIELD
    if FIELD is None:
        FIELD = ol.random_field(11, modes=8, kmax=3, std=0.5)
    return FIELD


def ops(hyg, gm, om, kind, minL, maxL):
    """GPU operator, FP64 oracle operator, FP32 oracle operator of one kind."""
    if kind == "const":
        return hyg.P1ConstantOperator(gm, minL, maxL), ol.Operator(om, minL, maxL, 8), ol.Operator(om, minL, maxL, 4)
    kf = ("fourier", 1.3, field())
    return (hyg.P1DivKGradOperator(gm, minL, maxL, 1.3, field()), ol.DivKGradOperator(om, minL, maxL, kf, 8),
            ol.DivKGradOperator(om, minL, maxL, kf, 4))


@pytest.fixture(scope="module")
def cube(hyg):
    hyg.init()
    return hyg.Mesh("cube_24el.msh"), ol.Mesh("cube_24el.msh")


@pytest.mark.parametrize("kind", ["const", "divkgrad"])
def test_power_iteration_matches_the_oracle(hyg, cube, kind):
    gm, om = cube
    L = 4
    n = gm.ndofs(L)
    v0 = rnd(n, 5)
    op, oop, _ = ops(hyg, gm, om, kind, L, L)
    x, tmp = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    x.upload(L, v0)
    got = op.power_iteration(x, tmp, 40, L)
    ov = ol.Function(om, L, L)
    ov.set(L, v0)
    ref = ol.lib().or_power_iteration(oop.h, ov.h, 40, L)
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)
    # the normalised iterate agrees too
    assert rel_inf(x.download(L), ov.get(L)) <= 1e-11


@pytest.mark.parametrize("kind,smoother", [("const", "jacobi"), ("const", "cheb"), ("divkgrad", "jacobi")])
def test_mixed_precision_refinement_matches_the_oracle(hyg, cube, kind, smoother):
    gm, om = cube
    minL, maxL = 0, 4
    n = gm.ndofs(maxL)
    _, bnd = gm.coords(maxL)
    b0 = rnd(n, 9)
    b0[bnd] = 0.0
    if smoother == "jacobi":
        kw = dict(pre=2, post=2, smoother="jacobi")
        ip = np.array([minL, maxL, 2, 2, 0, 0, 2, 10000], dtype=np.int32)
        dp = np.array([2.0 / 3.0, 0.0, 0.0, 1e-16])
    else:
        kw = dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * 1.9, 1.2 * 1.9))
        ip = np.array([minL, maxL, 3, 3, 0, 1, 2, 10000], dtype=np.int32)
        dp = np.array([2.0 / 3.0, 0.3 * 1.9, 1.2 * 1.9, 1e-16])
    op, oopd, oops = ops(hyg, gm, om, kind, minL, maxL)
    # FP64 reference solution from the oracle
    xd, bd = ol.Function(om, minL, maxL), ol.Function(om, minL, maxL)
    bd.set(maxL, b0)
    res = ol.gmg(oopd, xd, bd, minL, maxL, 60, **kw)
    assert res[-1] / res[0] < 1e-13
    xref = xd.get(maxL)
    # the oracle's defect correction: outer steps to 1e-10
    xo = ol.Function(om, minL, maxL)
    rir = np.zeros(41)
    ol.lib().or_ir(oopd.h, oops.h, xo.h, bd.h, ip, dp, 40, 1, rir)
    steps_ref = int(np.argmax(rir / rir[0] <= 1e-10))
    assert rir[steps_ref] / rir[0] <= 1e-10
    # the device
    x, b = hyg.Function(gm, minL, maxL), hyg.Function(gm, minL, maxL)
    b.upload(maxL, b0)
    g32 = hyg.GMG(op, minL, maxL, hyg.F32, **kw)
    got = hyg.ir_solve(op, g32, x, b, maxL, inner=1, max_outer=40, rel_tol=1e-10)
    assert got[-1] <= 1e-10 * got[0]
    assert abs(got[0] - rir[0]) <= 1e-12 * rir[0]
    assert abs((len(got) - 1) - steps_ref) <= 1, (len(got) - 1, steps_ref)
    assert rel_inf(x.download(maxL), xref) <= 5e-9


def _vcycle_graph_vs_oracle(hyg, hip_stream, gm, om, minL, maxL, p2=False):
    lat = 1 if p2 else 0
    n = gm.ndofs(maxL + lat)
    _, bnd = gm.coords(maxL + lat)
    b0 = rnd(n, 17)
    b0[bnd] = 0.0
    kw = dict(pre=2, post=2, smoother="jacobi", cg_tol=1e-10)
    sp = "p2" if p2 else "p1"
    op = hyg.P2ConstantOperator(gm, minL, maxL) if p2 else hyg.P1ConstantOperator(gm, minL, maxL)
    oop = ol.P2Operator(om, minL, maxL) if p2 else ol.Operator(om, minL, maxL)
    g = hyg.GMG(op, minL, maxL, **kw)
    b = hyg.Function(gm, minL, maxL, space=sp)
    b.upload(maxL, b0)
    x = hyg.Function(gm, minL, maxL, space=sp)
    ox, ob = ol.Function(om, minL + lat, maxL + lat), ol.Function(om, minL + lat, maxL + lat)
    ob.set(maxL + lat, b0)
    other, tmp = hyg.Function(gm, maxL, maxL, space=sp), hyg.Function(gm, maxL, maxL, space=sp)
    other.upload(maxL, b0)
    for k in range(3):
        g.vcycle_graph(x, b, maxL, 100, hip_stream)
        got = x.download(maxL)
        res = ol.gmg(oop, ox, ob, minL + lat, maxL + lat, 1, **kw)
        ref = ox.get(maxL + lat)
        assert rel_inf(got, ref) <= 1e-10, k
        if k == 0:
            assert res[1] < 0.5 * res[0]
        op.apply(other, tmp, maxL, hyg.INNER)
        _ = tmp.download(maxL)


def test_graph_vcycle_with_coarse_cg_work_cube24(hyg, hip_stream, cube):
    gm, om = cube
    _vcycle_graph_vs_oracle(hyg, hip_stream, gm, om, 1, 4)


def test_graph_vcycle_with_coarse_cg_work_cuboid(hyg, hip_stream):
    hyg.init()
    args = ((-1.0, 0.5, 2.0), (2.0, 1.5, 2.75), 3, 1, 2)
    gm, om = hyg.Mesh("symmetric_cuboid", *args), ol.Mesh("symmetric_cuboid", *args)
    _vcycle_graph_vs_oracle(hyg, hip_stream, gm, om, 0, 4)


def test_graph_vcycle_p2(hyg, hip_stream, cube):
    gm, om = cube
    _vcycle_graph_vs_oracle(hyg, hip_stream, gm, om, 0, 3, p2=True)


@pytest.mark.parametrize("smoother", ["sor", "ssor"])
def test_graph_vcycle_sor_matches_eager(hyg, hip_stream, cube, smoother):
    gm, _ = cube
    minL, maxL = 1, 5
    n = gm.ndofs(maxL)
    _, bnd = gm.coords(maxL)
    b0 = rnd(n, 23)
    b0[bnd] = 0.0
    op = hyg.P1ConstantOperator(gm, minL, maxL)
    g = hyg.GMG(op, minL, maxL, pre=2, post=2, smoother=smoother, relax=1.0, cg_tol=1e-10)
    b = hyg.Function(gm, minL, maxL)
    b.upload(maxL, b0)
    xe, xg = hyg.Function(gm, minL, maxL), hyg.Function(gm, minL, maxL)
    for k in range(3):
        g.vcycle(xe, b, maxL)
        g.vcycle_graph(xg, b, maxL, 100, hip_stream)
        ref = xe.download(maxL)
        assert rel_inf(xg.download(maxL), ref) <= 1e-12, k


def test_graph_is_recaptured_for_new_functions(hyg, hip_stream, cube):
    """The captured graph is keyed on the functions' identities, not on handle addresses: a function destroyed
    and replaced (possibly at the same address) gets a new capture on its own arrays."""
    gm, _ = cube
    minL, maxL = 1, 4
    n = gm.ndofs(maxL)
    _, bnd = gm.coords(maxL)
    b0 = rnd(n, 31)
    b0[bnd] = 0.0
    op = hyg.P1ConstantOperator(gm, minL, maxL)
    g = hyg.GMG(op, minL, maxL, pre=2, post=2, smoother="jacobi", cg_tol=1e-10)
    b = hyg.Function(gm, minL, maxL)
    b.upload(maxL, b0)
    xe = hyg.Function(gm, minL, maxL)
    g.vcycle(xe, b, maxL)
    ref = xe.download(maxL)
    for _ in range(3):
        x = hyg.Function(gm, minL, maxL)
        g.vcycle_graph(x, b, maxL, 100, hip_stream)
        assert rel_inf(x.download(maxL), ref) <= 1e-12
        del x


SHELL5 = ("shell", 5, 2, 0.55, 1.0)
COEF5 = dict(kind=0, k0=1.0, alpha=2.99, beta=0.1, seed=4321)


@pytest.fixture(scope="module")
def shell5(hyg):
    hyg.init()
    return hyg.Mesh(*SHELL5), ol.Mesh(*SHELL5)


@pytest.mark.parametrize("L", [1, 2])
def test_configuration5_operator_on_the_shell(hyg, shell5, L):
    gm, om = shell5
    n = gm.ndofs(L + 1)
    x0, d0 = rnd(n, 2), rnd(n, 4)
    op = hyg.P2DivKGradOperator(gm, L, L, **COEF5)
    oop = ol.P2DivKGradOperator(om, L, L, **COEF5)
    s, d = hyg.Function(gm, L, L, space="p2"), hyg.Function(gm, L, L, space="p2")
    os_, od = ol.Function(om, L + 1, L + 1), ol.Function(om, L + 1, L + 1)
    s.upload(L, x0)
    os_.set(L + 1, x0)
    for flag in (hyg.INNER, hyg.ALL):
        d.upload(L, d0)
        od.set(L + 1, d0)
        op.apply(s, d, L, flag, hyg.ADD)
        oop.apply(os_, od, L + 1, flag, True)
        assert rel_inf(d.download(L), od.get(L + 1)) <= TOL64, flag
    dinv = hyg.Function(gm, L, L, space="p2")
    op.inverse_diagonal(dinv, L)
    assert rel_inf(dinv.download(L), oop.inv_diag(L + 1)) <= TOL64


@pytest.mark.parametrize("stored", [False, True])
def test_configuration5_chebyshev_vcycle_history(hyg, shell5, stored):
    gm, om = shell5
    minL, maxL, cycles = 0, 2, 5
    n = gm.ndofs(maxL + 1)
    _, bnd = gm.coords(maxL + 1)
    b0 = rnd(n, 78)
    b0[bnd] = 0.0
    oop = ol.P2DivKGradOperator(om, minL, maxL, **COEF5)
    # Chebyshev bounds from the oracle's spectral radius estimate of D^-1 A (deterministic start vector)
    v = ol.Function(om, maxL + 1, maxL + 1)
    v.set(maxL + 1, rnd(n, 1))
    rho = ol.lib().or_power_iteration(oop.h, v.h, 30, maxL + 1)
    kw = dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * rho, 1.2 * rho))
    op = hyg.P2DivKGradOperator(gm, minL, maxL, **COEF5)
    if stored:
        op.store_local_matrices()
    x, b = hyg.Function(gm, minL, maxL, space="p2"), hyg.Function(gm, minL, maxL, space="p2")
    b.upload(maxL, b0)
    got = hyg.GMG(op, minL, maxL, **kw).solve(x, b, maxL, cycles)
    ox, ob = ol.Function(om, minL + 1, maxL + 1), ol.Function(om, minL + 1, maxL + 1)
    ob.set(maxL + 1, b0)
    ref = ol.gmg(oop, ox, ob, minL + 1, maxL + 1, cycles, **kw)
    assert ref[-1] < 0.5 * ref[0]
    assert np.all(np.abs(got - ref) <= 1e-10 * ref + 1e-14 * ref[0]), (got, ref)
    assert rel_inf(x.download(maxL), ox.get(maxL + 1)) <= 1e-10
