"""Configurations of the P2 cell-interior row kernel (hyg_op_tune / hyg_op_set_config on the P2 constant operator):
z-segments of 8, 16, 32 or 64 planes with one or two planes of loads in flight.

* Every configuration gives the first configuration's apply, apply-add, residual, Jacobi sweep and Chebyshev-smoothed
  V-cycle bit for bit (only the schedule changes, not the arithmetic), on the 24-tet cube at P2 levels 3 and 5, and the
  apply matches the oracle (1e-13 relative l-inf).
* hyg_op_tune returns a valid choice per mode, keeps it (hyg_op_get_config) and the tuned apply still matches.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


@pytest.fixture(scope="module")
def cube(hyg):
    hyg.init()
    return hyg.Mesh("cube_24el.msh"), ol.Mesh("cube_24el.msh")


def run_modes(hyg, gm, op, L, v):
    """results of apply, apply-add, residual, Jacobi sweep (modes 0..3) and of one V(1,1) Chebyshev cycle over
    levels L-1, L (modes 4 and 5 at level L)"""
    f = lambda lo=L: hyg.Function(gm, lo, L, space="p2")
    x, b, d = f(), f(), f()
    x.upload(L, v[0])
    b.upload(L, v[1])
    d.upload(L, v[2])
    out = []
    op.apply(x, d, L, hyg.INNER)
    out.append(d.download(L))
    d.upload(L, v[2])
    op.apply(x, d, L, hyg.INNER, hyg.ADD)
    out.append(d.download(L))
    d.upload(L, v[2])
    op.residual(d, b, x, L, hyg.INNER)
    out.append(d.download(L))
    d.upload(L, v[2])
    op.smooth_jac(d, b, x, 0.6, L, hyg.INNER)
    out.append(d.download(L))
    g = hyg.GMG(op, L - 1, L, pre=1, post=1, smoother="cheb", cheb_order=3, cheb_bounds=(0.5, 2.5))
    xs, bs = f(L - 1), f(L - 1)
    xs.upload(L, v[0])
    bs.upload(L, v[1])
    g.vcycle(xs, bs, L)
    out.append(xs.download(L))
    return out


@pytest.mark.parametrize("L", [3, 5])
def test_every_configuration_gives_the_same_results(hyg, cube, L):
    gm, om = cube
    op = hyg.P2ConstantOperator(gm, L - 1, L)
    n = gm.ndofs(L + 1)
    rng = np.random.default_rng(40 + L)
    v = [rng.uniform(-1, 1, n) for _ in range(3)]
    ref = None
    for cfg in range(8):
        for mode in range(6):
            op.set_config(L, hyg.F64, mode, cfg)
        assert op.config(L, hyg.F64, 0)[0] == cfg
        print(f"P2 row-kernel configuration {cfg}, P2 level {L}", flush=True)
        got = run_modes(hyg, gm, op, L, v)
        if ref is None:
            ref = got
            os_, od = ol.Function(om, L + 1, L + 1), ol.Function(om, L + 1, L + 1)
            os_.set(L + 1, v[0])
            od.set(L + 1, v[2])  # points outside the flag keep the destination's values, as on the device
            ol.P2Operator(om, L, L).apply(os_, od, L + 1, ol.INNER, False)
            assert rel_inf(got[0], od.get(L + 1)) <= 1e-13
        for k in range(5):
            assert np.array_equal(got[k], ref[k]), (cfg, k)
        # FP32 too
        for mode in range(6):
            op.set_config(L, hyg.F32, mode, cfg)
        s32, d32 = hyg.Function(gm, L, L, hyg.F32, space="p2"), hyg.Function(gm, L, L, hyg.F32, space="p2")
        s32.upload(L, v[0])
        d32.upload(L, v[2])
        op.apply(s32, d32, L, hyg.INNER)
        assert rel_inf(d32.download(L), ref[0]) <= 1e-5


def test_tuning_keeps_a_valid_choice(hyg, cube):
    gm, om = cube
    L = 5
    op = hyg.P2ConstantOperator(gm, L, L)
    print("tuning the P2 row kernel", flush=True)
    chosen = op.tune(L, hyg.F64, reps=3)
    assert len(chosen) == 6 and all(0 <= c < 8 for c in chosen), chosen
    for mode, c in enumerate(chosen):
        cfg, seg, depth = op.config(L, hyg.F64, mode)
        assert cfg == c and seg == (8, 16, 32, 64)[c // 2] and depth == c % 2 + 1
    n = gm.ndofs(L + 1)
    src = np.random.default_rng(8).uniform(-1, 1, n)
    s, d = hyg.Function(gm, L, L, space="p2"), hyg.Function(gm, L, L, space="p2")
    s.upload(L, src)
    op.apply(s, d, L, hyg.INNER)
    os_, od = ol.Function(om, L + 1, L + 1), ol.Function(om, L + 1, L + 1)
    os_.set(L + 1, src)
    ol.P2Operator(om, L, L).apply(os_, od, L + 1, ol.INNER, False)
    assert rel_inf(d.download(L), od.get(L + 1)) <= 1e-13
