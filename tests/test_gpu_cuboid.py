"""GPU side of MeshInfo::meshSymmetricCuboid (mesh/MeshGenSymmetricCuboid.cpp:35-253), MeshInfo::meshCuboid
(mesh/MeshGenCuboid.cpp:35-174) and MeshInfo::meshCubedDomain (mesh/MeshGenCubedDomain.cpp:37-383): the product's generators against the oracle's restatements (written
independently): identical primitive counts, DoF coordinates (bit for bit) and boundary flags; then the P1 apply
within 1e-13 relative l-inf and a V-cycle residual history within 1e-10 per cycle on those meshes."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
CASES = [("symmetric_cuboid", ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1, 1, 1)),
         ("symmetric_cuboid", ((-1.0, 0.5, 2.0), (2.0, 1.5, 2.75), 3, 1, 2)),
         ("cuboid", ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1, 1, 1)),
         ("cuboid", ((-1.0, 0.5, 2.0), (2.0, 1.5, 2.75), 3, 2, 2)),
         ("cubed_domain", ([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)], 0)),
         ("cubed_domain", ([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)], 1))]


def rel_inf(a, b):
    s = np.abs(b).max()
    return np.abs(np.asarray(a) - np.asarray(b)).max() / (s if s > 0 else 1.0)


@pytest.fixture(scope="module")
def meshes(hyg):
    hyg.init()
    return [(hyg.Mesh(kind, *c), ol.Mesh(kind, *c)) for kind, c in CASES]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_mesh_matches_the_oracle(hyg, meshes, k):
    gm, om = meshes[k]
    assert list(gm.info()) == list(om.info())
    for L in (0, 1, 3):
        gx, gb = gm.coords(L)
        ox, ob = om.coords(L)
        assert np.array_equal(gx, ox)
        assert np.array_equal(gb, ob)


@pytest.mark.parametrize("k", range(len(CASES)))
def test_p1_apply(hyg, meshes, k):
    gm, om = meshes[k]
    L = 4
    src = np.random.default_rng(3).uniform(-1.0, 1.0, gm.ndofs(L))
    op, oop = hyg.P1ConstantOperator(gm, L, L), ol.Operator(om, L, L)
    s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
    s.upload(L, src)
    os_.set(L, src)
    for flag in (hyg.INNER, hyg.ALL):
        op.apply(s, d, L, flag)
        od.set(L, np.zeros(gm.ndofs(L)))
        oop.apply(os_, od, L, flag, False)
        assert rel_inf(d.download(L), od.get(L)) <= 1e-13, flag


@pytest.mark.parametrize("k", [1, 3, 4, 5])
def test_vcycle_history(hyg, meshes, k):
    gm, om = meshes[k]
    minL, maxL, cycles = 0, 4, 5
    n = gm.ndofs(maxL)
    _, bnd = gm.coords(maxL)
    b0 = np.random.default_rng(77).uniform(-1.0, 1.0, n)
    b0[bnd] = 0.0
    kw = dict(pre=2, post=2, smoother="jacobi")
    op = hyg.P1ConstantOperator(gm, minL, maxL)
    x, b = hyg.Function(gm, minL, maxL), hyg.Function(gm, minL, maxL)
    b.upload(maxL, b0)
    got = hyg.GMG(op, minL, maxL, **kw).solve(x, b, maxL, cycles)
    oop = ol.Operator(om, minL, maxL)
    ox, ob = ol.Function(om, minL, maxL), ol.Function(om, minL, maxL)
    ob.set(maxL, b0)
    ref = ol.gmg(oop, ox, ob, minL, maxL, cycles, **kw)
    assert ref[-1] < 0.2 * ref[0]
    assert np.all(np.abs(got - ref) <= 1e-10 * ref + 1e-14 * ref[0]), (got, ref)


def test_p2_apply_and_transfers(hyg, meshes):
    gm, om = meshes[1]
    Lc = 2
    nc, nf = gm.ndofs(Lc + 1), gm.ndofs(Lc + 2)
    rng = np.random.default_rng(9)
    u, v = rng.uniform(-1, 1, nc), rng.uniform(-1, 1, nf)
    op, oop = hyg.P2ConstantOperator(gm, Lc + 1, Lc + 1), ol.P2Operator(om, Lc + 1, Lc + 1)
    s, d = hyg.Function(gm, Lc + 1, Lc + 1, space="p2"), hyg.Function(gm, Lc + 1, Lc + 1, space="p2")
    os_, od = ol.Function(om, Lc + 2, Lc + 2), ol.Function(om, Lc + 2, Lc + 2)
    s.upload(Lc + 1, v)
    os_.set(Lc + 2, v)
    op.apply(s, d, Lc + 1, hyg.ALL)
    oop.apply(os_, od, Lc + 2, ol.ALL, False)
    assert rel_inf(d.download(Lc + 1), od.get(Lc + 2)) <= 1e-13
    f = hyg.Function(gm, Lc, Lc + 1, space="p2")
    of = ol.Function(om, Lc + 1, Lc + 2)
    f.upload(Lc, u)
    of.set(Lc + 1, u)
    f.prolongate(Lc, hyg.ALL, hyg.REPLACE)
    ol.p2_prolongate(of, Lc + 1, hyg.ALL, False)
    assert rel_inf(f.download(Lc + 1), of.get(Lc + 2)) <= 1e-13
    f.upload(Lc + 1, v)
    of.set(Lc + 2, v)
    f.restrict(Lc + 1, hyg.ALL)
    ol.p2_restrict(of, Lc + 2, hyg.ALL)
    assert rel_inf(f.download(Lc), of.get(Lc + 1)) <= 1e-13
