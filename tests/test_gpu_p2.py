"""GPU parity of the P2 path (P2ConstantLaplaceOperator, quadratic prolongation and restriction, P2
multigrid) against the oracle, through the C ABI. A P2 function of level L holds the vertex lattice of level
L+1; oracle functions are created at those lattice levels. FP64 within 1e-13 relative l-inf for apply,
inverse diagonal, Jacobi sweep, residual and transfers; FP32 apply within 1e-5 of FP64; V-cycle residual
history within 1e-10 per cycle; A x = 0 on inner DoFs for linear x at full size."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
TOL64 = 1e-13
NAMES = ["cube_24el.msh", "cube_kuhn_6el.msh", "tet_1el.msh", "pyramid_tilted_4el.msh",
         "three_tets_with_two_joint_faces.msh"]


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


def rnd(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


@pytest.fixture(scope="module")
def meshes(hyg):
    hyg.init()
    return {n: (hyg.Mesh(n), ol.Mesh(n)) for n in NAMES}


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("L", [0, 1, 2, 3])
def test_apply_parity(hyg, meshes, name, L):
    gm, om = meshes[name]
    n = gm.ndofs(L + 1)
    src, old = rnd(n, 1), rnd(n, 2)
    op = hyg.P2ConstantOperator(gm, L, L)
    oop = ol.P2Operator(om, L, L)
    s, d = hyg.Function(gm, L, L, space="p2"), hyg.Function(gm, L, L, space="p2")
    os_, od = ol.Function(om, L + 1, L + 1), ol.Function(om, L + 1, L + 1)
    s.upload(L, src)
    os_.set(L + 1, src)
    for flag in (hyg.INNER, hyg.ALL):
        for upd in (hyg.REPLACE, hyg.ADD):
            d.upload(L, old)
            od.set(L + 1, old)
            op.apply(s, d, L, flag, upd)
            oop.apply(os_, od, L + 1, flag, upd == hyg.ADD)
            assert rel_inf(d.download(L), od.get(L + 1)) <= TOL64, (flag, upd)


def test_apply_parity_level5(hyg, meshes):
    gm, om = meshes["cube_24el.msh"]
    L = 5
    n = gm.ndofs(L + 1)
    src = rnd(n, 3)
    op = hyg.P2ConstantOperator(gm, L, L)
    oop = ol.P2Operator(om, L, L)
    s, d = hyg.Function(gm, L, L, space="p2"), hyg.Function(gm, L, L, space="p2")
    os_, od = ol.Function(om, L + 1, L + 1), ol.Function(om, L + 1, L + 1)
    s.upload(L, src)
    os_.set(L + 1, src)
    op.apply(s, d, L, hyg.INNER)
    oop.apply(os_, od, L + 1, ol.INNER)
    assert rel_inf(d.download(L), od.get(L + 1)) <= TOL64


@pytest.mark.parametrize("name", ["cube_24el.msh", "cube_kuhn_6el.msh", "three_tets_with_two_joint_faces.msh"])
@pytest.mark.parametrize("L", [1, 3])
def test_diagonal_smoother_residual(hyg, meshes, name, L):
    gm, om = meshes[name]
    n = gm.ndofs(L + 1)
    x0, b0, d0 = rnd(n, 4), rnd(n, 5), rnd(n, 6)
    op = hyg.P2ConstantOperator(gm, L, L)
    oop = ol.P2Operator(om, L, L)
    dinv = hyg.Function(gm, L, L, space="p2")
    op.inverse_diagonal(dinv, L)
    assert rel_inf(dinv.download(L), oop.inv_diag(L + 1)) <= TOL64
    x, b, d = (hyg.Function(gm, L, L, space="p2") for _ in range(3))
    x.upload(L, x0)
    b.upload(L, b0)
    d.upload(L, d0)
    op.smooth_jac(d, b, x, 0.6, L, hyg.INNER)
    ox, ob, od = (ol.Function(om, L + 1, L + 1) for _ in range(3))
    ox.set(L + 1, x0)
    ob.set(L + 1, b0)
    od.set(L + 1, d0)
    oop.smooth_jac(od, ob, ox, 0.6, L + 1, ol.INNER)
    assert rel_inf(d.download(L), od.get(L + 1)) <= TOL64
    r = hyg.Function(gm, L, L, space="p2")
    r.upload(L, d0)
    op.residual(r, b, x, L, hyg.INNER)
    t = ol.Function(om, L + 1, L + 1)
    oop.apply(ox, t, L + 1, ol.INNER, False)
    _, bnd = om.coords(L + 1)
    assert rel_inf(r.download(L), np.where(bnd, d0, b0 - t.get(L + 1))) <= TOL64


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("Lc", [0, 1, 2])
def test_transfer_parity(hyg, meshes, name, Lc):
    gm, om = meshes[name]
    nc, nf = gm.ndofs(Lc + 1), gm.ndofs(Lc + 2)
    u, old, v = rnd(nc, 7), rnd(nf, 8), rnd(nf, 9)
    f = hyg.Function(gm, Lc, Lc + 1, space="p2")
    of = ol.Function(om, Lc + 1, Lc + 2)
    for flag, upd in ((hyg.INNER | hyg.NEUMANN, hyg.REPLACE), (hyg.ALL, hyg.ADD), (hyg.INNER | hyg.NEUMANN, hyg.ADD),
                      (hyg.DIRICHLET, hyg.REPLACE), (hyg.DIRICHLET, hyg.ADD)):
        f.upload(Lc, u)
        f.upload(Lc + 1, old)
        of.set(Lc + 1, u)
        of.set(Lc + 2, old)
        f.prolongate(Lc, flag, upd)
        ol.p2_prolongate(of, Lc + 1, flag, upd == hyg.ADD)
        assert rel_inf(f.download(Lc + 1), of.get(Lc + 2)) <= TOL64, (flag, upd)
    # DIRICHLET alone: boundary points and (as in the reference, whose cell loop takes no flag) cell interiors
    for flag in (hyg.ALL, hyg.INNER | hyg.NEUMANN, hyg.DIRICHLET):
        f.upload(Lc + 1, v)
        f.upload(Lc, u)
        of.set(Lc + 2, v)
        of.set(Lc + 1, u)
        f.restrict(Lc + 1, flag)
        ol.p2_restrict(of, Lc + 2, flag)
        assert rel_inf(f.download(Lc), of.get(Lc + 1)) <= TOL64, flag


@pytest.mark.parametrize("name,smoother", [("cube_24el.msh", "jacobi"), ("cube_kuhn_6el.msh", "jacobi"),
                                           ("cube_24el.msh", "cheb")])
def test_vcycle_history_parity(hyg, meshes, name, smoother):
    gm, om = meshes[name]
    minL, maxL, cycles = 0, 3, 5
    n = gm.ndofs(maxL + 1)
    _, bnd = gm.coords(maxL + 1)
    b0 = rnd(n, 77)
    b0[bnd] = 0.0
    kw = dict(pre=2, post=2, smoother="jacobi") if smoother == "jacobi" else \
        dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * 2.2, 1.2 * 2.2))
    op = hyg.P2ConstantOperator(gm, minL, maxL)
    x, b = hyg.Function(gm, minL, maxL, space="p2"), hyg.Function(gm, minL, maxL, space="p2")
    b.upload(maxL, b0)
    got = hyg.GMG(op, minL, maxL, **kw).solve(x, b, maxL, cycles)
    oop = ol.P2Operator(om, minL, maxL)
    ox, ob = ol.Function(om, minL + 1, maxL + 1), ol.Function(om, minL + 1, maxL + 1)
    ob.set(maxL + 1, b0)
    ref = ol.gmg(oop, ox, ob, minL + 1, maxL + 1, cycles, **kw)
    assert ref[-1] < 0.05 * ref[0]
    assert np.all(np.abs(got - ref) <= 1e-10 * ref + 1e-14 * ref[0]), (got, ref)
    assert rel_inf(x.download(maxL), ox.get(maxL + 1)) <= 1e-10


def test_fp32_apply_and_mixed_solve(hyg, meshes):
    gm, _ = meshes["cube_24el.msh"]
    L = 4
    n = gm.ndofs(L + 1)
    x0 = rnd(n, 10)
    op = hyg.P2ConstantOperator(gm, 0, L)
    s32, y32 = hyg.Function(gm, L, L, hyg.F32, space="p2"), hyg.Function(gm, L, L, hyg.F32, space="p2")
    s64, y64 = hyg.Function(gm, L, L, space="p2"), hyg.Function(gm, L, L, space="p2")
    s32.upload(L, x0)
    s64.upload(L, x0)
    op.apply(s32, y32, L)
    op.apply(s64, y64, L)
    assert rel_inf(y32.download(L), y64.download(L)) <= 1e-5
    # FP32 V-cycles inside FP64 defect correction reach 1e-10
    _, bnd = gm.coords(L + 1)
    b0 = rnd(n, 11)
    b0[bnd] = 0.0
    b, x = hyg.Function(gm, 0, L, space="p2"), hyg.Function(gm, 0, L, space="p2")
    b.upload(L, b0)
    g32 = hyg.GMG(op, 0, L, hyg.F32, pre=2, post=2, smoother="jacobi")
    res = hyg.ir_solve(op, g32, x, b, L, inner=1, max_outer=40, rel_tol=1e-10)
    assert res[-1] <= 1e-10 * res[0]


def test_linear_functions_at_full_size(hyg, meshes):
    """A x = 0 on inner DoFs for a linear x, P2 level 6 on the 24-tet cube (lattice level 7, 8.5M DoFs)."""
    gm, _ = meshes["cube_24el.msh"]
    L = 6
    xyz, bnd = gm.coords(L + 1)
    op = hyg.P2ConstantOperator(gm, L, L)
    s, d = hyg.Function(gm, L, L, space="p2"), hyg.Function(gm, L, L, space="p2")
    s.upload(L, 1.0 + 2.0 * xyz[:, 0] - xyz[:, 1] + 0.5 * xyz[:, 2])
    op.apply(s, d, L, hyg.INNER)
    assert np.abs(d.download(L)[~bnd]).max() < 1e-9


def test_p1_and_p2_handles_do_not_mix(hyg, meshes):
    gm, _ = meshes["cube_kuhn_6el.msh"]
    op2 = hyg.P2ConstantOperator(gm, 1, 1)
    p1a, p1b = hyg.Function(gm, 1, 1), hyg.Function(gm, 1, 1)
    with pytest.raises(hyg.HygError):
        op2.apply(p1a, p1b, 1)
    p2 = hyg.Function(gm, 1, 1, space="p2")
    with pytest.raises(hyg.HygError):
        p2.assign([1.0], [p1a], 1)
