"""GPU parity of the spherical shell mesh and of the P2 variable-coefficient operator (P2ElementwiseOperator with
p2_div_k_grad_affine_q4) against the oracle, through the C ABI: the shell's DoF coordinates and boundary flags
agree exactly; FP64 apply, inverse diagonal and Jacobi sweep within 1e-13 relative l-inf; FP32 apply within 1e-5;
V-cycle residual history on the shell within 1e-10 per cycle; the same with stored per-point stencils."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
TOL64 = 1e-13
SHELL = ("shell", 3, 2, 0.55, 1.0)
COEF = dict(kind=0, k0=1.0, alpha=2.99, beta=0.1, seed=1234)


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


def rnd(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


@pytest.fixture(scope="module")
def meshes(hyg):
    hyg.init()
    return {"shell": (hyg.Mesh(*SHELL), ol.Mesh(*SHELL)),
            "shell5": (hyg.Mesh("shell", 5, 2, 0.55, 1.0), ol.Mesh("shell", 5, 2, 0.55, 1.0)),
            "cube_24el.msh": (hyg.Mesh("cube_24el.msh"), ol.Mesh("cube_24el.msh"))}


@pytest.mark.parametrize("name", ["shell", "shell5"])
def test_shell_mesh_matches_the_oracle(hyg, meshes, name):
    gm, om = meshes[name]
    assert list(gm.info()) == list(om.info())
    for L in (0, 1, 3):
        gx, gb = gm.coords(L)
        ox, ob = om.coords(L)
        assert np.abs(gx - ox).max() <= 1e-15 * np.abs(ox).max()
        assert np.array_equal(gb, ob)


def test_p1_apply_on_the_shell(hyg, meshes):
    gm, om = meshes["shell5"]
    L = 3
    n = gm.ndofs(L)
    src = rnd(n, 1)
    op, oop = hyg.P1ConstantOperator(gm, L, L), ol.Operator(om, L, L)
    s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
    s.upload(L, src)
    os_.set(L, src)
    op.apply(s, d, L, hyg.INNER)
    oop.apply(os_, od, L, ol.INNER, False)
    assert rel_inf(d.download(L), od.get(L)) <= TOL64


@pytest.mark.parametrize("name,L", [("cube_24el.msh", 0), ("cube_24el.msh", 1), ("cube_24el.msh", 3), ("shell", 0), ("shell", 1),
                                    ("shell", 2)])
@pytest.mark.parametrize("coef", ["exp", "poly"])
@pytest.mark.parametrize("stored", [False, True])
def test_apply_diagonal_smoother_parity(hyg, meshes, name, L, coef, stored):
    """On the fly, and with the per-point stencils stored (computeAndStoreLocalElementMatrices)."""
    gm, om = meshes[name]
    kw = COEF if coef == "exp" else dict(kind=1)
    n = gm.ndofs(L + 1)
    x0, b0, d0 = rnd(n, 2), rnd(n, 3), rnd(n, 4)
    op = hyg.P2DivKGradOperator(gm, L, L, **kw)
    if stored:
        op.store_local_matrices()
    oop = ol.P2DivKGradOperator(om, L, L, **kw)
    s, d = hyg.Function(gm, L, L, space="p2"), hyg.Function(gm, L, L, space="p2")
    os_, od = ol.Function(om, L + 1, L + 1), ol.Function(om, L + 1, L + 1)
    s.upload(L, x0)
    os_.set(L + 1, x0)
    for flag in (hyg.INNER, hyg.ALL):
        for upd in (hyg.REPLACE, hyg.ADD):
            d.upload(L, d0)
            od.set(L + 1, d0)
            op.apply(s, d, L, flag, upd)
            oop.apply(os_, od, L + 1, flag, upd == hyg.ADD)
            assert rel_inf(d.download(L), od.get(L + 1)) <= TOL64, (flag, upd)
    dinv = hyg.Function(gm, L, L, space="p2")
    op.inverse_diagonal(dinv, L)
    assert rel_inf(dinv.download(L), oop.inv_diag(L + 1)) <= TOL64
    b = hyg.Function(gm, L, L, space="p2")
    b.upload(L, b0)
    d.upload(L, d0)
    op.smooth_jac(d, b, s, 0.6, L, hyg.INNER)
    ob = ol.Function(om, L + 1, L + 1)
    ob.set(L + 1, b0)
    od.set(L + 1, d0)
    oop.smooth_jac(od, ob, os_, 0.6, L + 1, ol.INNER)
    assert rel_inf(d.download(L), od.get(L + 1)) <= TOL64
    # FP32 apply against FP64
    s32, y32 = hyg.Function(gm, L, L, hyg.F32, space="p2"), hyg.Function(gm, L, L, hyg.F32, space="p2")
    s32.upload(L, x0)
    op.apply(s32, y32, L, hyg.ALL)
    op.apply(s, d, L, hyg.ALL)
    assert rel_inf(y32.download(L), d.download(L)) <= 1e-5


@pytest.mark.parametrize("stored", [False, True])
def test_vcycle_history_on_the_shell(hyg, meshes, stored):
    gm, om = meshes["shell"]
    minL, maxL, cycles = 0, 2, 5
    n = gm.ndofs(maxL + 1)
    _, bnd = gm.coords(maxL + 1)
    b0 = rnd(n, 77)
    b0[bnd] = 0.0
    kw = dict(pre=2, post=2, smoother="jacobi")
    op = hyg.P2DivKGradOperator(gm, minL, maxL, **COEF)
    if stored:
        op.store_local_matrices()
    x, b = hyg.Function(gm, minL, maxL, space="p2"), hyg.Function(gm, minL, maxL, space="p2")
    b.upload(maxL, b0)
    got = hyg.GMG(op, minL, maxL, **kw).solve(x, b, maxL, cycles)
    oop = ol.P2DivKGradOperator(om, minL, maxL, **COEF)
    ox, ob = ol.Function(om, minL + 1, maxL + 1), ol.Function(om, minL + 1, maxL + 1)
    ob.set(maxL + 1, b0)
    ref = ol.gmg(oop, ox, ob, minL + 1, maxL + 1, cycles, **kw)
    assert ref[-1] < 0.2 * ref[0]
    assert np.all(np.abs(got - ref) <= 1e-10 * ref + 1e-14 * ref[0]), (got, ref)
    assert rel_inf(x.download(maxL), ox.get(maxL + 1)) <= 1e-10


@pytest.mark.parametrize("name,L", [("shell", 2), ("shell5", 2)])
def test_p2_constant_apply_and_smoother_on_the_shell(hyg, meshes, name, L):
    """The constant P2 operator's stencil kernels (cell interiors, one-sided face stencils, edges and vertices)
    on macro-cells of every orientation the shell produces."""
    gm, om = meshes[name]
    n = gm.ndofs(L + 1)
    x0, b0, d0 = rnd(n, 12), rnd(n, 13), rnd(n, 14)
    op, oop = hyg.P2ConstantOperator(gm, L, L), ol.P2Operator(om, L, L)
    s, d = hyg.Function(gm, L, L, space="p2"), hyg.Function(gm, L, L, space="p2")
    os_, od = ol.Function(om, L + 1, L + 1), ol.Function(om, L + 1, L + 1)
    s.upload(L, x0)
    os_.set(L + 1, x0)
    for flag in (hyg.INNER, hyg.ALL):
        for upd in (hyg.REPLACE, hyg.ADD):
            d.upload(L, d0)
            od.set(L + 1, d0)
            op.apply(s, d, L, flag, upd)
            oop.apply(os_, od, L + 1, flag, upd == hyg.ADD)
            assert rel_inf(d.download(L), od.get(L + 1)) <= TOL64, (flag, upd)
    b = hyg.Function(gm, L, L, space="p2")
    b.upload(L, b0)
    d.upload(L, d0)
    op.smooth_jac(d, b, s, 0.6, L, hyg.INNER)
    ob = ol.Function(om, L + 1, L + 1)
    ob.set(L + 1, b0)
    od.set(L + 1, d0)
    oop.smooth_jac(od, ob, os_, 0.6, L + 1, ol.INNER)
    assert rel_inf(d.download(L), od.get(L + 1)) <= TOL64


@pytest.mark.parametrize("stored", [False, True])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_vector_apply_matches_component_applies(hyg, meshes, stored, dtype):
    """VectorLaplaceOperator::apply through hyg_apply_vector: bit-identical to applying the scalar operator to each
    component (the fused stored pass computes every component's sums in the same order)."""
    gm, _ = meshes["shell"]
    L = 2
    dt = hyg.F64 if dtype == "f64" else hyg.F32
    n = gm.ndofs(L + 1)
    op = hyg.P2DivKGradOperator(gm, L, L, **COEF)
    if stored:
        op.store_local_matrices()
    xs = [hyg.Function(gm, L, L, dt, space="p2") for _ in range(3)]
    ys = [hyg.Function(gm, L, L, dt, space="p2") for _ in range(3)]
    zs = [hyg.Function(gm, L, L, dt, space="p2") for _ in range(3)]
    for k in range(3):
        xs[k].upload(L, rnd(n, 40 + k))
        ys[k].upload(L, rnd(n, 50 + k))
        zs[k].upload(L, rnd(n, 50 + k))
    for flag, upd in ((hyg.INNER, hyg.REPLACE), (hyg.ALL, hyg.ADD)):
        hyg.apply_vector(op, xs, ys, L, flag, upd)
        for k in range(3):
            op.apply(xs[k], zs[k], L, flag, upd)
            assert np.array_equal(ys[k].download(L), zs[k].download(L)), (flag, upd, k)
