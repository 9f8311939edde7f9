"""GPU parity of the variable-coefficient operator -div(k grad u) (P1ElementwiseOperator with
p1_div_k_grad_affine_q3) against the oracle: FP64 within 1e-13 relative l-inf for apply, inverse
diagonal, Jacobi sweep and residual; FP32 apply within 1e-5 of FP64; V-cycle residual history within
1e-10 per cycle."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
TOL64 = 1e-13
MODES = ol.random_field(11, modes=8, kmax=3)


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


def rnd(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


@pytest.fixture(scope="module")
def meshes(hyg):
    hyg.init()
    names = ["cube_24el.msh", "cube_kuhn_6el.msh", "pyramid_tilted_4el.msh", "three_tets_with_two_joint_faces.msh"]
    return {n: (hyg.Mesh(n), ol.Mesh(n)) for n in names}


@pytest.mark.parametrize("name", ["cube_24el.msh", "cube_kuhn_6el.msh", "pyramid_tilted_4el.msh",
                                  "three_tets_with_two_joint_faces.msh"])
@pytest.mark.parametrize("L", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("field", ["const", "fourier"])
def test_apply_and_diagonal_parity(hyg, meshes, name, L, field):
    gm, om = meshes[name]
    modes = np.zeros((0, 5)) if field == "const" else MODES
    k0 = 2.5 if field == "const" else 1.3
    op = hyg.P1DivKGradOperator(gm, L, L, k0, modes)
    oop = ol.DivKGradOperator(om, L, L, ("fourier", k0, modes))
    n = gm.ndofs(L)
    src, d0 = rnd(n, 50 + L), rnd(n, 60 + L)
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
            assert rel_inf(d.download(L), od.get(L)) <= TOL64, (name, L, field, flag, update)
    dinv = hyg.Function(gm, L, L)
    op.inverse_diagonal(dinv, L)
    assert rel_inf(dinv.download(L), oop.inv_diag(L)) <= TOL64


@pytest.mark.parametrize("L", [3, 5])
@pytest.mark.parametrize("stored", [False, True])
def test_smoother_residual_and_fp32(hyg, meshes, L, stored):
    """On the fly, and with the coefficient integrals kept (computeAndStoreLocalElementMatrices)."""
    gm, om = meshes["cube_24el.msh"]
    n = gm.ndofs(L)
    x0, b0, d0 = rnd(n, 1), rnd(n, 2), rnd(n, 3)
    op = hyg.P1DivKGradOperator(gm, L, L, 1.3, MODES)
    if stored:
        op.store_local_matrices()
    oop = ol.DivKGradOperator(om, L, L, ("fourier", 1.3, MODES))
    x, b, d = (hyg.Function(gm, L, L) for _ in range(3))
    x.upload(L, x0)
    b.upload(L, b0)
    d.upload(L, d0)
    op.smooth_jac(d, b, x, 0.6, L, hyg.INNER)
    ox, ob, od = (ol.Function(om, L, L) for _ in range(3))
    ox.set(L, x0)
    ob.set(L, b0)
    od.set(L, d0)
    oop.smooth_jac(od, ob, ox, 0.6, L, ol.INNER)
    assert rel_inf(d.download(L), od.get(L)) <= TOL64
    r = hyg.Function(gm, L, L)
    r.upload(L, d0)
    op.residual(r, b, x, L, hyg.INNER)
    oo = ol.Function(om, L, L)
    oo.set(L, d0)
    oop.apply(ox, oo, L, ol.INNER, False)
    _, bnd = om.coords(L)
    assert rel_inf(r.download(L), np.where(bnd, d0, b0 - oo.get(L))) <= TOL64
    # FP32 apply against the FP64 result
    s32, y32 = hyg.Function(gm, L, L, hyg.F32), hyg.Function(gm, L, L, hyg.F32)
    s32.upload(L, x0)
    op.apply(s32, y32, L, hyg.INNER)
    y64 = hyg.Function(gm, L, L)
    op.apply(x, y64, L, hyg.INNER)
    assert rel_inf(y32.download(L), y64.download(L)) <= 1e-5


@pytest.mark.parametrize("name,stored", [("cube_24el.msh", False), ("cube_kuhn_6el.msh", False), ("cube_24el.msh", True)])
def test_vcycle_history_parity(hyg, meshes, name, stored):
    gm, om = meshes[name]
    minL, maxL, cycles = 0, 4, 5
    n = gm.ndofs(maxL)
    _, bnd = gm.coords(maxL)
    b0 = rnd(n, 77)
    b0[bnd] = 0.0
    kw = dict(pre=2, post=2, smoother="jacobi")
    op = hyg.P1DivKGradOperator(gm, minL, maxL, 1.3, MODES)
    if stored:
        op.store_local_matrices()
    x, b = hyg.Function(gm, minL, maxL), hyg.Function(gm, minL, maxL)
    b.upload(maxL, b0)
    got = hyg.GMG(op, minL, maxL, **kw).solve(x, b, maxL, cycles)
    oop = ol.DivKGradOperator(om, minL, maxL, ("fourier", 1.3, MODES))
    ox, ob = ol.Function(om, minL, maxL), ol.Function(om, minL, maxL)
    ob.set(maxL, b0)
    ref = ol.gmg(oop, ox, ob, minL, maxL, cycles, **kw)
    assert np.all(np.abs(got - ref) <= 1e-10 * ref + 1e-14 * ref[0]), (got, ref)
    assert rel_inf(x.download(maxL), ox.get(maxL)) <= 1e-10


def test_full_coefficient_field_level6(hyg, meshes):
    """32 modes of wavenumber <= 8 (the variable-coefficient workload's field) on the 24-tet cube."""
    gm, om = meshes["cube_24el.msh"]
    L = 6
    modes = ol.random_field(2024, modes=32, kmax=8)
    n = gm.ndofs(L)
    src = rnd(n, 5)
    op = hyg.P1DivKGradOperator(gm, L, L, 1.0, modes)
    s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    s.upload(L, src)
    op.apply(s, d, L, hyg.INNER)
    oop = ol.DivKGradOperator(om, L, L, ("fourier", 1.0, modes))
    os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
    os_.set(L, src)
    oop.apply(os_, od, L, ol.INNER, False)
    assert rel_inf(d.download(L), od.get(L)) <= TOL64
