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


def _sym_error(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case", ["p1_cheb", "p1_sor_ordered", "p1_sor", "p2_cheb"])
def test_graph_vcycle_with_coarse_cg_work(hyg, hip_stream, case):
    """The coarse grid has inner DoFs (24-tet cube: 13 at P1 level 1; at P2 level 0 the interior edge midpoints
    and the centre), so the captured coarse CG (device-side step lengths, two-slot ping-pong, tolerance gating)
    takes real steps. Tolerance 0 and a fixed iteration count make the eager CG, the captured CG and the
    oracle's CGSolver take the same iterations. Every replay equals the eager cycle and the oracle's V-cycle."""
    import oracle_lib as ol
    hyg.init()
    gm, om = hyg.Mesh("cube_24el.msh"), ol.Mesh("cube_24el.msh")
    p2 = case.startswith("p2")
    minL, maxL = (0, 3) if p2 else (1, 5)
    shift = 1 if p2 else 0
    space = "p2" if p2 else "p1"
    n = gm.ndofs(maxL + shift)
    _, bnd = gm.coords(maxL + shift)
    b0 = np.random.default_rng(17).uniform(-1, 1, n)
    b0[bnd] = 0.0
    cg_iters = 16
    smoother = case.split("_", 1)[1]
    kw = dict(pre=2, post=2, smoother=smoother, cg_max_iter=cg_iters, cg_tol=0.0)
    if smoother == "cheb":
        rho = 2.2 if p2 else 1.9
        kw.update(pre=3, post=3, cheb_order=2, cheb_bounds=(0.3 * rho, 1.2 * rho))
    if smoother == "sor":
        kw.update(relax=1.2)
    if smoother == "sor_ordered":
        kw.update(relax=1.0)
    op = hyg.P2ConstantOperator(gm, minL, maxL) if p2 else hyg.P1ConstantOperator(gm, minL, maxL)
    g = hyg.GMG(op, minL, maxL, **kw)
    f = lambda: hyg.Function(gm, minL, maxL, space=space)
    b = f()
    b.upload(maxL, b0)
    xe, xg = f(), f()
    if smoother != "sor":  # the multi-colour SOR has no oracle counterpart
        oop = ol.P2Operator(om, minL, maxL) if p2 else ol.Operator(om, minL, maxL)
        ox, ob = ol.Function(om, minL + shift, maxL + shift), ol.Function(om, minL + shift, maxL + shift)
        ob.set(maxL + shift, b0)
    other, tmp = f(), f()
    other.upload(maxL, b0)
    for k in range(3):
        print(f"graph V-cycle with coarse CG work: {case}, cycle {k}", flush=True)
        g.vcycle(xe, b, maxL)
        eager = xe.download(maxL)
        g.vcycle_graph(xg, b, maxL, cg_iters, hip_stream)
        got = xg.download(maxL)
        op.apply(other, tmp, maxL, hyg.INNER)
        _ = tmp.download(maxL)
        assert _sym_error(got, eager) <= 1e-12, (case, k)
        if smoother != "sor":
            okw = {key: v for key, v in kw.items()}
            ol.gmg(oop, ox, ob, minL + shift, maxL + shift, 1, **okw)
            assert _sym_error(got, ox.get(maxL + shift)) <= 1e-10, (case, k)
    # the cycles converge, so the coarse solves did their work
    r = f()
    op.residual(r, b, xg, maxL, hyg.INNER)
    assert np.abs(r.download(maxL)).max() < 0.2 * np.abs(b0).max()
