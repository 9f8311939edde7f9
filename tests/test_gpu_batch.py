"""Batched independent problems (hyg_batch_*): every grid operation of all problems in one launch.

A batch mixes the 24-tet and the Kuhn 6-tet cube, levels 2..5, the constant Laplacian and -div(k grad u) with a
seeded Fourier-mode coefficient (local matrices stored). Checked per problem:
  * the batched FP64 apply against the oracle (1e-13 relative l-inf), with the per-level tuned and with forced
    cell-kernel configurations;
  * batched FP64 Chebyshev V(3,3) and Jacobi V(2,2) cycles: residual history against the oracle's GMG (1e-10
    relative per cycle), and the iterate bit for bit equal to the unbatched hyg_gmg_vcycle (the batched launch runs
    the same kernel bodies with the same decomposition);
  * the captured batched cycle equal to the eager one on every replay, with eager work between replays;
  * batched mixed-precision defect correction: every problem reaches 1e-10, the residual histories equal those of
    the unbatched hyg_ir_solve bit for bit, and the solution agrees with the FP64 oracle solve.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
MODES = ol.random_field(21, modes=8, kmax=3)
CHEB = dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * 1.9, 1.2 * 1.9))
JAC = dict(pre=2, post=2, smoother="jacobi")
# (mesh, level, coefficient)
PROBLEMS = [("cube_24el.msh", 4, "const"), ("cube_kuhn_6el.msh", 5, "fourier"), ("cube_24el.msh", 2, "fourier"),
            ("cube_kuhn_6el.msh", 3, "const"), ("cube_24el.msh", 3, "fourier"), ("cube_kuhn_6el.msh", 4, "const")]


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


@pytest.fixture(scope="module")
def meshes(hyg):
    hyg.init()
    return {n: (hyg.Mesh(n), ol.Mesh(n)) for n in ("cube_24el.msh", "cube_kuhn_6el.msh")}


def make_ops(hyg, meshes, minL=0):
    ops, oops = [], []
    for name, L, field in PROBLEMS:
        gm, om = meshes[name]
        if field == "const":
            ops.append(hyg.P1ConstantOperator(gm, minL, L))
            oops.append(ol.Operator(om, minL, L))
        else:
            op = hyg.P1DivKGradOperator(gm, minL, L, 1.3, MODES)
            op.store_local_matrices()
            ops.append(op)
            oops.append(ol.DivKGradOperator(om, minL, L, ("fourier", 1.3, MODES)))
    return ops, oops


def rhs(gm, L, seed):
    n = gm.ndofs(L)
    _, bnd = gm.coords(L)
    b0 = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    b0[bnd] = 0.0
    return b0


@pytest.mark.parametrize("cfg", [-1, 0, 3, 6])
def test_batched_apply_matches_the_oracle(hyg, meshes, cfg):
    ops, oops = make_ops(hyg, meshes)
    gs, xs, bs, srcs, dsts, data = [], [], [], [], [], []
    for k, ((name, L, _), op) in enumerate(zip(PROBLEMS, ops)):
        gm, _ = meshes[name]
        gs.append(hyg.GMG(op, 0, L, **CHEB))
        xs.append(hyg.Function(gm, 0, L))
        bs.append(hyg.Function(gm, 0, L))
        s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
        v = np.random.default_rng(500 + k).uniform(-1.0, 1.0, gm.ndofs(L))
        s.upload(L, v)
        srcs.append(s)
        dsts.append(d)
        data.append(v)
    bt = hyg.Batch(gs, xs, bs)
    bt.set_config(cfg)
    for flag in (hyg.INNER, hyg.ALL):
        print(f"batched apply, configuration {cfg}, flag {flag}", flush=True)
        bt.apply(srcs, dsts, flag)
        for k, ((name, L, field), oop) in enumerate(zip(PROBLEMS, oops)):
            _, om = meshes[name]
            os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
            os_.set(L, data[k])
            oop.apply(os_, od, L, flag, False)
            assert rel_inf(dsts[k].download(L), od.get(L)) <= 1e-13, (name, L, field, flag)


@pytest.mark.parametrize("kw", [CHEB, JAC], ids=["cheb", "jacobi"])
def test_batched_vcycles_match_the_oracle_and_the_unbatched_cycle(hyg, meshes, kw):
    ops, oops = make_ops(hyg, meshes)
    cycles = 3
    gs, xs, bs, b0s = [], [], [], []
    for k, ((name, L, _), op) in enumerate(zip(PROBLEMS, ops)):
        gm, _ = meshes[name]
        gs.append(hyg.GMG(op, 0, L, **kw))
        x, b = hyg.Function(gm, 0, L), hyg.Function(gm, 0, L)
        b0 = rhs(gm, L, 700 + k)
        b.upload(L, b0)
        xs.append(x)
        bs.append(b)
        b0s.append(b0)
    bt = hyg.Batch(gs, xs, bs)
    hist = [bt.residual_norms()]
    for c in range(cycles):
        print(f"batched V-cycle {c}", flush=True)
        bt.vcycle()
        hist.append(bt.residual_norms())
    hist = np.array(hist).T
    for k, ((name, L, field), op, oop) in enumerate(zip(PROBLEMS, ops, oops)):
        gm, om = meshes[name]
        ox, ob = ol.Function(om, 0, L), ol.Function(om, 0, L)
        ob.set(L, b0s[k])
        ref = ol.gmg(oop, ox, ob, 0, L, cycles, **kw)
        assert np.all(np.abs(hist[k] - ref) / ref <= 1e-10), (name, L, field, hist[k], ref)
        # the unbatched cycles give the same iterate bit for bit
        g1 = hyg.GMG(op, 0, L, **kw)
        x1, b1 = hyg.Function(gm, 0, L), hyg.Function(gm, 0, L)
        b1.upload(L, b0s[k])
        for _ in range(cycles):
            g1.vcycle(x1, b1, L)
        assert np.array_equal(xs[k].download(L), x1.download(L)), (name, L, field)


def test_captured_batched_cycle_equals_eager(hyg, meshes, hip_stream):
    ops, _ = make_ops(hyg, meshes)
    runs = {}
    for mode in ("eager", "graph"):
        gs, xs, bs = [], [], []
        for k, ((name, L, _), op) in enumerate(zip(PROBLEMS, ops)):
            gm, _ = meshes[name]
            gs.append(hyg.GMG(op, 0, L, **CHEB))
            x, b = hyg.Function(gm, 0, L), hyg.Function(gm, 0, L)
            b.upload(L, rhs(gm, L, 900 + k))
            xs.append(x)
            bs.append(b)
        bt = hyg.Batch(gs, xs, bs)
        out = []
        for c in range(4):
            print(f"{mode} batched cycle {c}", flush=True)
            bt.vcycle(use_graph=(mode == "graph"), stream=hip_stream)
            out.append([x.download(L) for x, (_, L, _) in zip(xs, PROBLEMS)])  # eager work between replays
        runs[mode] = out
    for c in range(4):
        for k in range(len(PROBLEMS)):
            assert np.array_equal(runs["graph"][c][k], runs["eager"][c][k]), (c, PROBLEMS[k])


@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_batched_defect_correction(hyg, meshes, hip_stream, use_graph):
    ops, oops = make_ops(hyg, meshes)
    gs32, x32, b32, x64, b64, b0s = [], [], [], [], [], []
    for k, ((name, L, _), op) in enumerate(zip(PROBLEMS, ops)):
        gm, _ = meshes[name]
        gs32.append(hyg.GMG(op, 0, L, hyg.F32, **CHEB))
        x32.append(hyg.Function(gm, 0, L, hyg.F32))
        b32.append(hyg.Function(gm, 0, L, hyg.F32))
        x, b = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
        b0 = rhs(gm, L, 1100 + k)
        b.upload(L, b0)
        x64.append(x)
        b64.append(b)
        b0s.append(b0)
    bt = hyg.Batch(gs32, x32, b32)
    print(f"batched defect correction, graph {use_graph}", flush=True)
    hist = bt.ir_solve(x64, b64, inner=1, max_outer=40, rel_tol=1e-10, use_graph=use_graph, stream=hip_stream)
    assert np.all(hist[:, -1] <= 1e-10 * hist[:, 0]), hist[:, -1] / hist[:, 0]
    for k, ((name, L, field), op, oop) in enumerate(zip(PROBLEMS, ops, oops)):
        gm, om = meshes[name]
        # unbatched defect correction: the same residual history, bit for bit, while this problem iterates
        g1 = hyg.GMG(op, 0, L, hyg.F32, **CHEB)
        x1, b1 = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
        b1.upload(L, b0s[k])
        h1 = hyg.ir_solve(op, g1, x1, b1, L, inner=1, max_outer=40, rel_tol=1e-10)
        m = len(h1)
        assert np.array_equal(hist[k, :m], h1), (name, L, field)
        # the solution agrees with the FP64 oracle solve (30 FP64 V-cycles)
        ox, ob = ol.Function(om, 0, L), ol.Function(om, 0, L)
        ob.set(L, b0s[k])
        ol.gmg(oop, ox, ob, 0, L, 30, **CHEB)
        ref = ox.get(L)
        assert rel_inf(x64[k].download(L), ref) <= 5e-9, (name, L, field)


def test_batch_rejects_incompatible_solvers(hyg, meshes):
    gm, _ = meshes["cube_24el.msh"]
    op = hyg.P1ConstantOperator(gm, 0, 3)
    x, b = hyg.Function(gm, 0, 3), hyg.Function(gm, 0, 3)
    ga = hyg.GMG(op, 0, 3, **CHEB)
    gb = hyg.GMG(op, 0, 3, **JAC)
    with pytest.raises(hyg.HygError) as e:
        hyg.Batch([ga, gb], [x, x], [b, b])
    assert e.value.code == hyg.ERR_ARG
    opk = hyg.P1DivKGradOperator(gm, 0, 3, 1.0, MODES)  # local matrices not stored
    gk = hyg.GMG(opk, 0, 3, **CHEB)
    with pytest.raises(hyg.HygError) as e:
        hyg.Batch([gk], [x], [b])
    assert e.value.code == hyg.ERR_ARG
    # a coarse level with inner DoFs (level 1 of the 24-tet cube has 13)
    g1 = hyg.GMG(op, 1, 3, **CHEB)
    x1, b1 = hyg.Function(gm, 1, 3), hyg.Function(gm, 1, 3)
    with pytest.raises(hyg.HygError) as e:
        hyg.Batch([g1], [x1], [b1])
    assert e.value.code == hyg.ERR_ARG
