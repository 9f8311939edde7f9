"""Reference-layout halos on the device (hyg_halo_*) against the oracle.

  * hyg_halo_pack of a function equals the oracle's memory with ghost layers (pinned geometrically on the CPU,
    tests/test_oracle_halo.py) element for element, FP64, on four meshes and levels 0..4;
  * unpack round trip: the own values unpacked into a zero function reproduce every point of macro-vertices, -edges
    and -faces and leave cell interiors at zero; shifted own values land in every replica (the apply of the result
    matches the oracle's apply of the shifted canonical vector);
  * hyg_halo_sum_replicas: a replica set becomes the sum of its cells' values: n(p) v(p) for a consistent function,
    n(p) = number of cells containing the point's primitive; primitives outside the flag are unchanged.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
MESHES = ["cube_24el.msh", "cube_kuhn_6el.msh", "pyramid_tilted_4el.msh", "three_tets_with_two_joint_faces.msh"]


@pytest.fixture(scope="module")
def meshes(hyg):
    hyg.init()
    return {n: (hyg.Mesh(n), ol.Mesh(n)) for n in MESHES}


def lower_dim_mask(om, L):
    """Canonical DoFs on macro-vertices, -edges and -faces, and the number of cells containing each one's primitive."""
    nv, ne, nf, nc = (int(v) for v in om.info()[:4])
    N = 2 ** L
    edges, faces = om.prims()
    cells = [set(c) for c in om.cells()]
    cnt = [sum(1 for c in cells if v in c) for v in range(nv)]
    for e in edges:
        cnt += [sum(1 for c in cells if set(e) <= c)] * max(0, N - 1)
    for t in faces:
        cnt += [sum(1 for c in cells if set(t) <= c)] * max(0, (N - 1) * (N - 2) // 2)
    n = om.ndofs(L)
    mask = np.zeros(n, dtype=bool)
    mask[: len(cnt)] = True
    counts = np.ones(n)
    counts[: len(cnt)] = cnt
    return mask, counts


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("L", [0, 2, 4])
def test_pack_matches_the_oracle(hyg, meshes, name, L):
    gm, om = meshes[name]
    n = gm.ndofs(L)
    v = np.random.default_rng(31 + L).uniform(-1, 1, n)
    f = hyg.Function(gm, L, L)
    f.upload(L, v)
    print(f"halo pack {name} L{L}", flush=True)
    got = hyg.halo_pack(f, L)
    off = hyg.halo_layout(gm, L)
    assert len(got) == off[-1]
    of = ol.Function(om, L, L)
    of.set(L, v)
    ref = np.zeros(len(got))
    ol.lib().or_halo_pack(of.h, L, ref)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("name", MESHES)
def test_unpack_round_trip_and_replicas(hyg, meshes, name):
    gm, om = meshes[name]
    L = 3
    n = gm.ndofs(L)
    v = np.random.default_rng(7).uniform(-1, 1, n)
    f = hyg.Function(gm, L, L)
    f.upload(L, v)
    mem = hyg.halo_pack(f, L)
    g = hyg.Function(gm, L, L)
    print(f"halo unpack {name}", flush=True)
    hyg.halo_unpack(g, L, mem)
    mask, _ = lower_dim_mask(om, L)
    got = g.download(L)
    assert np.array_equal(got[mask], v[mask])
    assert np.all(got[~mask] == 0.0)
    # shifted own values: every replica must take them (the apply sees consistent replicas)
    off = hyg.halo_layout(gm, L)
    shifted = mem.copy()
    for k in range(len(off) - 1):
        shifted[off[k]:off[k + 1]] += 1.0
    hyg.halo_unpack(f, L, shifted)
    want = v + mask * 1.0
    assert np.array_equal(f.download(L), want)
    op = hyg.P1ConstantOperator(gm, L, L)
    d = hyg.Function(gm, L, L)
    op.apply(f, d, L, hyg.ALL)
    oop = ol.Operator(om, L, L)
    os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
    os_.set(L, want)
    oop.apply(os_, od, L, ol.ALL, False)
    ref = od.get(L)
    assert np.abs(d.download(L) - ref).max() <= 1e-13 * np.abs(ref).max()


def test_unpack_follows_the_downward_chain(hyg, meshes):
    """Own values that disagree at shared points: the vertex value wins over the edge's end, the edge over the face's
    border (VertexDoFPackInfo's vertex -> edge -> face order)."""
    gm, om = meshes["cube_24el.msh"]
    L = 2
    n = gm.ndofs(L)
    f = hyg.Function(gm, L, L)
    f.upload(L, np.zeros(n))
    off = hyg.halo_layout(gm, L)
    nv, ne = (int(v) for v in gm.info()[:2])
    mem = np.zeros(off[-1])
    for k in range(nv):
        mem[off[k]] = 1.0  # every vertex own value
    for k in range(nv, nv + ne):
        mem[off[k]:off[k] + 2 ** L + 1] = 2.0  # every edge's own values, ends included
    for k in range(nv + ne, len(off) - 1):
        w = 2 ** L + 1
        mem[off[k]:off[k] + w * (w + 1) // 2] = 3.0  # every face's own values, border included
    hyg.halo_unpack(f, L, mem)
    N = 2 ** L
    got = f.download(L)
    nf = int(gm.info()[2])
    want = np.concatenate([np.full(nv, 1.0), np.full(ne * (N - 1), 2.0), np.full(nf * (N - 1) * (N - 2) // 2, 3.0)])
    assert np.array_equal(got[: len(want)], want)
    assert np.all(got[len(want):] == 0.0)


@pytest.mark.parametrize("name", MESHES)
def test_sum_replicas(hyg, meshes, name):
    gm, om = meshes[name]
    L = 3
    n = gm.ndofs(L)
    v = np.random.default_rng(8).uniform(-1, 1, n)
    mask, counts = lower_dim_mask(om, L)
    f = hyg.Function(gm, L, L)
    f.upload(L, v)
    print(f"halo sum {name}", flush=True)
    hyg.halo_sum_replicas(f, L, hyg.ALL)
    assert np.abs(f.download(L) - counts * v).max() <= 1e-15 * np.abs(counts * v).max() * 4
    # only Inner primitives: Dirichlet points keep their value
    f.upload(L, v)
    hyg.halo_sum_replicas(f, L, hyg.INNER)
    _, bnd = gm.coords(L)
    want = np.where(bnd, v, counts * v)
    assert np.abs(f.download(L) - want).max() <= 1e-15 * np.abs(want).max() * 4
