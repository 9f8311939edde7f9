"""A V-cycle captured as a hipGraph and replayed gives the eager V-cycle's result, on every replay,
with other eager work and synchronisations between the replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,smoother", [("cube_24el.msh", "cheb"), ("cube_kuhn_6el.msh", "jacobi")])
@pytest.mark.parametrize("op_kind", ["const", "divkgrad"])
def test_graph_vcycle_matches_eager(hyg, hip_stream, name, smoother, op_kind):
    hyg.init()
    gm = hyg.Mesh(name)
    minL, maxL = 0, 5
    n = gm.ndofs(maxL)
    _, bnd = gm.coords(maxL)
    b0 = np.random.default_rng(7).uniform(-1, 1, n)
    b0[bnd] = 0.0
    if op_kind == "const":
        op = hyg.P1ConstantOperator(gm, minL, maxL)
    else:
        import oracle_lib as ol
        op = hyg.P1DivKGradOperator(gm, minL, maxL, 1.3, ol.random_field(3, modes=8, kmax=3))
    kw = dict(pre=2, post=2, smoother="jacobi") if smoother == "jacobi" else \
        dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * 1.9, 1.2 * 1.9))
    g = hyg.GMG(op, minL, maxL, **kw)
    b = hyg.Function(gm, minL, maxL)
    b.upload(maxL, b0)
    # eager reference: the iterate after each of three cycles
    xe = hyg.Function(gm, minL, maxL)
    ref = []
    for _ in range(3):
        g.vcycle(xe, b, maxL)
        ref.append(xe.download(maxL))
    # graph: captured on the first call, replayed afterwards; unrelated eager work in between
    xg = hyg.Function(gm, minL, maxL)
    other, tmp = hyg.Function(gm, maxL, maxL), hyg.Function(gm, maxL, maxL)
    other.upload(maxL, b0)
    for k in range(3):
        g.vcycle_graph(xg, b, maxL, 8, hip_stream)
        got = xg.download(maxL)
        op.apply(other, tmp, maxL, hyg.INNER)
        _ = tmp.download(maxL)
        scale = np.abs(ref[k]).max()
        assert np.abs(got - ref[k]).max() <= 1e-12 * scale, (name, op_kind, k)
