"""Multi-colour SOR on the device (hyg_smooth_sor; SORSmoothable::smooth_sor, P1Operator.hpp:343-423) against the
restatement in sor_model.py driven by the oracle's operator: the vertex colouring, forward and backward sweeps in
FP64 (<= 1e-13 relative) and FP32, a size-independent Gauss-Seidel property at level 7, and GMG with the SOR
smoother (SORSmoother.hpp:38-50)."""
import numpy as np
import pytest

import oracle_lib as ol
import sor_model as sm
from conftest import load_hyg

pytestmark = pytest.mark.gpu
hyg = load_hyg()


@pytest.fixture(autouse=True, scope="module")
def _init():
    hyg.init()

MESHES = ["cube_24el.msh", "cube_6el.msh", "regular_octahedron_8el.msh", "tet_1el.msh"]


def _colouring(name):
    om = ol.Mesh(name)
    X = om.coords(0)[0]
    cells = om.cells()
    m, vcol = sm.vertex_colouring(len(X), cells)
    return om, X, cells, m, vcol


@pytest.mark.parametrize("name", MESHES)
def test_vertex_colouring_matches(name):
    om, X, cells, m, vcol = _colouring(name)
    gm = hyg.Mesh(name)
    gmc, gvcol = gm.sor_colours()
    assert gmc == m
    assert np.array_equal(gvcol, vcol)


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("L", [2, 3, 5])
@pytest.mark.parametrize("relax,backwards", [(1.0, False), (1.3, False), (1.3, True)])
def test_sor_sweep_f64(name, L, relax, backwards):
    om, X, cells, m, vcol = _colouring(name)
    gm = hyg.Mesh(name)
    xyz, bnd = om.coords(L)
    col = sm.point_colours(xyz, X, cells, vcol, m, 1 << L)
    rng = np.random.default_rng(L * 31 + int(backwards))
    x0 = rng.uniform(-1, 1, len(col))
    b = rng.uniform(-1, 1, len(col))
    # device sweep
    op = hyg.P1ConstantOperator(gm, L, L)
    x, bf = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    x.upload(L, x0)
    bf.upload(L, b)
    print(f"sor {name} L{L} relax {relax} backwards {backwards}", flush=True)
    op.smooth_sor(x, bf, relax, L, hyg.INNER, backwards)
    got = x.download(L)
    # restatement with the oracle's operator
    oop = ol.Operator(om, L, L)
    src, dst = ol.Function(om, L, L), ol.Function(om, L, L)

    def apply(v):
        src.set(L, v)
        oop.apply(src, dst, L, ol.ALL, False)
        return dst.get(L)

    ref = sm.sor_sweep(apply, oop.inv_diag(L), x0, b, relax, col, m, ~bnd, backwards)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.array_equal(got[bnd], x0[bnd])


def test_sor_sweep_f32():
    name, L = "cube_24el.msh", 5
    gm = hyg.Mesh(name)
    n = gm.ndofs(L)
    rng = np.random.default_rng(11)
    x0, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    out = {}
    for dt in (hyg.F64, hyg.F32):
        op = hyg.P1ConstantOperator(gm, L, L)
        x, bf = hyg.Function(gm, L, L, dt), hyg.Function(gm, L, L, dt)
        x.upload(L, x0)
        bf.upload(L, b)
        op.smooth_sor(x, bf, 1.2, L)
        out[dt] = x.download(L)
    assert np.abs(out[hyg.F32] - out[hyg.F64]).max() <= 2e-5 * np.abs(out[hyg.F64]).max()


def test_last_colour_residual_zero_level7():
    # relax = 1: after a forward sweep the residual vanishes on the inner points of the last colour, and the
    # energy J(x) = x.Ax / 2 - b.x (zero boundary values) decreases
    name, L = "cube_24el.msh", 7
    om, X, cells, m, vcol = _colouring(name)
    gm = hyg.Mesh(name)
    xyz, bnd = gm.coords(L)
    col = sm.point_colours(xyz, X, cells, vcol, m, 1 << L)
    n = len(col)
    rng = np.random.default_rng(5)
    x0, b0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    x0[bnd] = 0.0
    b0[bnd] = 0.0
    op = hyg.P1ConstantOperator(gm, L, L)
    x, bf, r = hyg.Function(gm, L, L), hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    x.upload(L, x0)
    bf.upload(L, b0)

    def energy():
        xv = x.download(L)
        op.apply(x, r, L, hyg.INNER)
        return 0.5 * xv @ r.download(L) - b0 @ xv

    e0 = energy()
    op.residual(r, bf, x, L)
    r0 = np.abs(r.download(L)).max()
    op.smooth_sor(x, bf, 1.0, L)
    op.residual(r, bf, x, L)
    rv = r.download(L)
    sel = (~bnd) & (col == m - 1)
    assert sel.sum() > 0
    assert np.abs(rv[sel]).max() <= 1e-12 * r0
    assert energy() < e0


def test_gmg_with_sor_smoother_beats_jacobi():
    name, L = "cube_24el.msh", 6
    gm = hyg.Mesh(name)
    n = gm.ndofs(L)
    _, bnd = gm.coords(L)
    bv = np.random.default_rng(2).uniform(-1, 1, n)
    bv[bnd] = 0.0
    op = hyg.P1ConstantOperator(gm, 0, L)
    b = hyg.Function(gm, 0, L)
    b.upload(L, bv)
    rate = {}
    for sm_, relax in (("sor", 1.0), ("ssor", 1.0), ("jacobi", 2.0 / 3.0)):
        x = hyg.Function(gm, 0, L)
        res = hyg.GMG(op, 0, L, pre=2, post=2, smoother=sm_, relax=relax).solve(x, b, L, 8)
        rate[sm_] = (res[-1] / res[-5]) ** 0.25
    print("asymptotic rates", rate)
    assert rate["sor"] < 0.35 and rate["sor"] < rate["jacobi"], rate
    # a symmetric sweep does a forward and a backward sweep: at least as good as the forward one
    assert rate["ssor"] <= rate["sor"], rate


def test_gmg_gauss_seidel_octahedron_reference_bound():
    # P1GMG3DConvergenceTest.cpp:36-125: regular_octahedron_8el.msh, levels 0..3, Gauss-Seidel V(3,3), u random in
    # the interior and sin(x) sinh(y) z on the boundary, f = 0; every cycle must reduce the squared discrete L2
    # residual by a factor below 3.2e-2
    name, minL, L = "regular_octahedron_8el.msh", 0, 3
    gm = hyg.Mesh(name)
    xyz, bnd = gm.coords(L)
    u0 = np.random.default_rng(0).uniform(0.0, 1.0, len(xyz))
    u0[bnd] = np.sin(xyz[bnd, 0]) * np.sinh(xyz[bnd, 1]) * xyz[bnd, 2]
    op = hyg.P1ConstantOperator(gm, minL, L)
    x, b = hyg.Function(gm, minL, L), hyg.Function(gm, minL, L)
    x.upload(L, u0)
    res = hyg.GMG(op, minL, L, pre=3, post=3, smoother="sor", relax=1.0).solve(x, b, L, 4)
    sq = (np.array(res[1:]) / np.array(res[:-1])) ** 2
    print("squared residual ratios", sq)
    assert np.all(sq < 3.2e-2), sq
