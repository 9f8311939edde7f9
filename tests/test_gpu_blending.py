"""GPU parity of the spherical-shell blending (hyg_mesh_set_shell_map: IcosahedralShellMap::setMap) and of the blended
P1 Laplacian (hyg_op_create_p1_blending_laplace: P1ElementwiseBlendingLaplaceOperator, p1_diffusion_blending_q3)
against the oracle (pinned on the CPU in tests/test_oracle_shell_map.py), through the C ABI:

* blended DoF coordinates equal the oracle's (1e-14);
* FP64 apply (Inner / All, replace / add), inverse diagonal and Jacobi sweep within 1e-13 relative l-inf on three
  shells, levels 1..3 (and 4 on one); FP32 apply within 1e-5;
* Jacobi V(2,2) cycles on the blended shell: residual history within 1e-10 per cycle;
* argument errors: no map, a mesh that is not a shell, surrogates of the blended operator.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
SHELLS = [(3, 2, 0.55, 1.0), (5, 2, 0.55, 1.0), (3, 3, 0.5, 1.0)]


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


@pytest.fixture(scope="module")
def shells(hyg):
    hyg.init()
    out = {}
    for a in SHELLS:
        gm, om = hyg.Mesh("shell", *a), ol.Mesh("shell", *a)
        gm.set_shell_map()
        om.set_shell_map()
        out[a] = (gm, om)
    return out


@pytest.mark.parametrize("args", SHELLS)
@pytest.mark.parametrize("L", [1, 2, 3])
def test_blended_operator_matches_the_oracle(hyg, shells, args, L):
    gm, om = shells[args]
    gx, gb = gm.coords(L)
    ox, ob = om.coords(L)
    assert np.abs(gx - ox).max() <= 1e-14 and np.array_equal(gb, ob)
    print(f"blended Laplacian shell {args} level {L}", flush=True)
    op = hyg.P1BlendingLaplaceOperator(gm, L, L)
    oop = ol.BlendingLaplaceOperator(om, L, L)
    n = gm.ndofs(L)
    rng = np.random.default_rng(60 + L)
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
            assert rel_inf(d.download(L), od.get(L)) <= 1e-13, (flag, update)
    dinv = hyg.Function(gm, L, L)
    op.inverse_diagonal(dinv, L)
    assert rel_inf(dinv.download(L), oop.inv_diag(L)) <= 1e-13
    s, b, out = hyg.Function(gm, L, L), hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    s.upload(L, src)
    b.upload(L, rhs)
    out.upload(L, src)
    op.smooth_jac(out, b, s, 0.6, L, hyg.INNER)
    oo, orr = ol.Function(om, L, L), ol.Function(om, L, L)
    oo.set(L, src)
    orr.set(L, rhs)
    oop.smooth_jac(oo, orr, os_, 0.6, L, ol.INNER)
    assert rel_inf(out.download(L), oo.get(L)) <= 1e-13
    s32, d32 = hyg.Function(gm, L, L, hyg.F32), hyg.Function(gm, L, L, hyg.F32)
    s32.upload(L, src)
    op.apply(s32, d32, L, hyg.ALL)
    od.set(L, np.zeros(n))
    oop.apply(os_, od, L, ol.ALL, False)
    assert rel_inf(d32.download(L), od.get(L)) <= 1e-5


def test_level4_apply(hyg, shells):
    gm, om = shells[(3, 2, 0.55, 1.0)]
    L = 4
    op = hyg.P1BlendingLaplaceOperator(gm, L, L)
    oop = ol.BlendingLaplaceOperator(om, L, L)
    src = np.random.default_rng(9).uniform(-1, 1, gm.ndofs(L))
    s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    s.upload(L, src)
    op.apply(s, d, L, hyg.INNER)
    os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
    os_.set(L, src)
    oop.apply(os_, od, L, ol.INNER, False)
    assert rel_inf(d.download(L), od.get(L)) <= 1e-13


def test_blended_vcycles(hyg, shells):
    gm, om = shells[(3, 2, 0.55, 1.0)]
    L, cycles = 3, 4
    op = hyg.P1BlendingLaplaceOperator(gm, 0, L)
    _, bnd = gm.coords(L)
    b0 = np.random.default_rng(12).uniform(-1, 1, len(bnd))
    b0[bnd] = 0.0
    x, b = hyg.Function(gm, 0, L), hyg.Function(gm, 0, L)
    b.upload(L, b0)
    print("blended V-cycles", flush=True)
    res = hyg.GMG(op, 0, L, pre=2, post=2, smoother="jacobi").solve(x, b, L, cycles)
    ox, ob = ol.Function(om, 0, L), ol.Function(om, 0, L)
    ob.set(L, b0)
    ref = ol.gmg(ol.BlendingLaplaceOperator(om, 0, L), ox, ob, 0, L, cycles, pre=2, post=2)
    assert np.all(np.abs(res - ref) <= 1e-10 * ref), (res, ref)
    assert res[-1] < 0.2 * res[0]


def test_blending_argument_errors(hyg):
    hyg.init()
    plain = hyg.Mesh("shell", 3, 2, 0.55, 1.0)
    with pytest.raises(hyg.HygError) as e:
        hyg.P1BlendingLaplaceOperator(plain, 1, 1)
    assert e.value.code == hyg.ERR_ARG
    cube = hyg.Mesh("cube_24el.msh")
    with pytest.raises(hyg.HygError) as e:
        cube.set_shell_map()
    assert e.value.code == hyg.ERR_ARG
    plain.set_shell_map()
    op = hyg.P1BlendingLaplaceOperator(plain, 2, 3)
    with pytest.raises(hyg.HygError) as e:
        hyg.check(hyg.load().hyg_op_interpolate_stencils(op.h, 2, 3))
    assert e.value.code == hyg.ERR_ARG
