"""GPU parity of SOR in the reference's order (hyg_smooth_sor_ordered: SORSmoothable::smooth_sor,
P1Operator.hpp:343-416) against the oracle's restatement (pinned on the CPU against a triangular solve and a
global-matrix model, tests/test_oracle_sor_ordered.py), through the C ABI.

* FP64 sweeps forward and backward, flags Inner and All, relax 1.3, five meshes, levels 1..5: 1e-12 relative l-inf
  (the sums run in another order; the Gauss-Seidel chain carries the rounding along); FP32 within 1e-5.
* Three sweeps in a row (the tables and snapshot are reused) against three oracle sweeps.
* At level 6 (no oracle): with relax = 1 and flag Inner, the residual vanishes at the lexicographically last interior
  point of every cell (it is the cell's last update and all its neighbours are final by then).
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
MESHES = ["cube_24el.msh", "cube_kuhn_6el.msh", "regular_octahedron_8el.msh", "three_tets_with_two_joint_faces.msh",
          "tet_tilted_1el.msh"]


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


@pytest.fixture(scope="module")
def meshes(hyg):
    hyg.init()
    return {n: (hyg.Mesh(n), ol.Mesh(n)) for n in MESHES}


def oracle_sweeps(om, L, x0, b, relax, flag, backwards, n=1):
    op = ol.Operator(om, L, L)
    x, r = ol.Function(om, L, L), ol.Function(om, L, L)
    x.set(L, x0)
    r.set(L, b)
    for _ in range(n):
        assert ol.lib().or_smooth_sor_ordered(op.h, x.h, r.h, relax, L, flag, int(backwards)) == 0
    return x.get(L)


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_ordered_sor_matches_the_oracle(hyg, meshes, name, L):
    gm, om = meshes[name]
    op = hyg.P1ConstantOperator(gm, L, L)
    n = gm.ndofs(L)
    rng = np.random.default_rng(100 + L)
    x0, b0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    for flag in (hyg.INNER, hyg.ALL):
        for backwards in (False, True):
            print(f"ordered SOR {name} L{L} flag {flag} backwards {backwards}", flush=True)
            x, b = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
            x.upload(L, x0)
            b.upload(L, b0)
            op.smooth_sor_ordered(x, b, 1.3, L, flag, backwards)
            ref = oracle_sweeps(om, L, x0, b0, 1.3, flag, backwards)
            assert rel_inf(x.download(L), ref) <= 1e-12, (flag, backwards)
            x32, b32 = hyg.Function(gm, L, L, hyg.F32), hyg.Function(gm, L, L, hyg.F32)
            x32.upload(L, x0)
            b32.upload(L, b0)
            op.smooth_sor_ordered(x32, b32, 1.3, L, flag, backwards)
            assert rel_inf(x32.download(L), ref) <= 1e-5, (flag, backwards)


def test_repeated_sweeps(hyg, meshes):
    gm, om = meshes["cube_24el.msh"]
    L = 4
    op = hyg.P1ConstantOperator(gm, L, L)
    n = gm.ndofs(L)
    rng = np.random.default_rng(7)
    x0, b0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    x, b = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    x.upload(L, x0)
    b.upload(L, b0)
    for _ in range(3):
        op.smooth_sor_ordered(x, b, 1.1, L, hyg.INNER)
    assert rel_inf(x.download(L), oracle_sweeps(om, L, x0, b0, 1.1, ol.INNER, False, 3)) <= 1e-12


def test_level6_invariants(hyg, meshes):
    gm, _ = meshes["cube_24el.msh"]
    L, N = 6, 64
    op = hyg.P1ConstantOperator(gm, L, L)
    n = gm.ndofs(L)
    _, bnd = gm.coords(L)
    rng = np.random.default_rng(21)
    x0, b0 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    x0[bnd] = 0.0
    b0[bnd] = 0.0
    x, b = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    x.upload(L, x0)
    b.upload(L, b0)
    print("ordered SOR level 6", flush=True)
    op.smooth_sor_ordered(x, b, 1.0, L, hyg.INNER)
    nv, ne, nf, nc = (int(v) for v in gm.info()[:4])
    nI = (N - 1) * (N - 2) * (N - 3) // 6
    first = nv + ne * (N - 1) + nf * (N - 1) * (N - 2) // 2
    last = [first + (c + 1) * nI - 1 for c in range(nc)]
    # residual b - A x at those points, from an apply
    ax = hyg.Function(gm, L, L)
    op.apply(x, ax, L, hyg.INNER)
    res = b0 - ax.download(L)
    assert np.abs(res[last]).max() <= 1e-12 * np.abs(b0).max()


def test_reference_convergence_test_in_reference_order(hyg):
    # P1GMG3DConvergenceTest.cpp:36-125 with the reference's own Gauss-Seidel order: regular_octahedron_8el.msh,
    # levels 0..3, V(3,3), u random in the interior and sin(x) sinh(y) z on the boundary, f = 0; every cycle reduces
    # the squared residual by a factor below 3.2e-2, and the residual history is the oracle's (1e-10 per cycle)
    name, minL, L, cycles = "regular_octahedron_8el.msh", 0, 3, 4
    gm, om = hyg.Mesh(name), ol.Mesh(name)
    xyz, bnd = gm.coords(L)
    u0 = np.random.default_rng(0).uniform(0.0, 1.0, len(xyz))
    u0[bnd] = np.sin(xyz[bnd, 0]) * np.sinh(xyz[bnd, 1]) * xyz[bnd, 2]
    op = hyg.P1ConstantOperator(gm, minL, L)
    x, b = hyg.Function(gm, minL, L), hyg.Function(gm, minL, L)
    x.upload(L, u0)
    print("GMG with ordered Gauss-Seidel", flush=True)
    res = hyg.GMG(op, minL, L, pre=3, post=3, smoother="sor_ordered", relax=1.0).solve(x, b, L, cycles)
    sq = (np.array(res[1:]) / np.array(res[:-1])) ** 2
    assert np.all(sq < 3.2e-2), sq
    ox, ob = ol.Function(om, minL, L), ol.Function(om, minL, L)
    ox.set(L, u0)
    ref = ol.gmg(ol.Operator(om, minL, L), ox, ob, minL, L, cycles, pre=3, post=3, smoother="sor_ordered", relax=1.0)
    assert np.all(np.abs(res - ref) <= 1e-10 * ref), (res, ref)


@pytest.mark.parametrize("name", ["cube_24el.msh", "cube_kuhn_6el.msh"])
def test_symmetric_ordered_sor_vcycles_match_the_oracle(hyg, name):
    minL, L, cycles = 0, 4, 3
    gm, om = hyg.Mesh(name), ol.Mesh(name)
    _, bnd = gm.coords(L)
    b0 = np.random.default_rng(33).uniform(-1, 1, len(bnd))
    b0[bnd] = 0.0
    op = hyg.P1ConstantOperator(gm, minL, L)
    x, b = hyg.Function(gm, minL, L), hyg.Function(gm, minL, L)
    b.upload(L, b0)
    print(f"GMG with symmetric ordered SOR on {name}", flush=True)
    res = hyg.GMG(op, minL, L, pre=2, post=2, smoother="ssor_ordered", relax=1.2).solve(x, b, L, cycles)
    ox, ob = ol.Function(om, minL, L), ol.Function(om, minL, L)
    ob.set(L, b0)
    ref = ol.gmg(ol.Operator(om, minL, L), ox, ob, minL, L, cycles, pre=2, post=2, smoother="ssor_ordered", relax=1.2)
    assert np.all(np.abs(res - ref) <= 1e-10 * ref), (res, ref)
