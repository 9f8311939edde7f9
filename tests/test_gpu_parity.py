"""Parity of the HIP library (through its C ABI) with the CPU oracle on the same inputs.

FP64 results must match the oracle within 1e-13 relative l-infinity (apply, smoother, residual,
transfer) and per-cycle residual norms within 1e-10 relative (V-cycles).
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

MESHES = ["cube_24el.msh", "cube_kuhn_6el.msh", "tet_1el.msh", "cube_6el.msh"]
TOL64 = 1e-13


def rel_inf(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).max()
    s = np.abs(np.asarray(b)).max()
    return d / s if s > 0 else d


def rnd(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


@pytest.fixture(scope="module")
def meshes(hyg):
    hyg.init()
    return {name: (hyg.Mesh(name), ol.Mesh(name)) for name in MESHES}


@pytest.mark.parametrize("name", MESHES)
def test_mesh_and_dof_order(hyg, meshes, name):
    gm, om = meshes[name]
    assert (gm.info() == om.info()).all()
    for L in (0, 1, 2, 3, 4):
        assert gm.ndofs(L) == om.ndofs(L)
        gx, gb = gm.coords(L)
        ox, ob = om.coords(L)
        assert np.array_equal(gb, ob)
        assert np.abs(gx - ox).max() < 1e-14


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("L", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("flag", ["inner", "all"])
@pytest.mark.parametrize("update", ["replace", "add"])
def test_apply_parity(hyg, meshes, name, L, flag, update):
    gm, om = meshes[name]
    n = gm.ndofs(L)
    src, dst0 = rnd(n, 11 + L), rnd(n, 97 + L)
    fl = hyg.INNER if flag == "inner" else hyg.ALL
    # GPU
    op = hyg.P1ConstantOperator(gm, L, L)
    s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    s.upload(L, src)
    d.upload(L, dst0)
    op.apply(s, d, L, fl, hyg.ADD if update == "add" else hyg.REPLACE)
    got = d.download(L)
    # oracle
    oop = ol.Operator(om, L, L)
    os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
    os_.set(L, src)
    od.set(L, dst0)
    oop.apply(os_, od, L, fl, update == "add")
    ref = od.get(L)
    assert rel_inf(got, ref) <= TOL64, (name, L, flag, update)
    # stencils agree entry by entry
    if L >= 2:
        for c in range(int(gm.info()[3])):
            assert rel_inf(op.cell_stencil(L, c), oop.cell_stencil(L, c)) <= 1e-14


@pytest.mark.parametrize("name", ["cube_24el.msh", "cube_kuhn_6el.msh"])
@pytest.mark.parametrize("L", [2, 3, 4])
def test_smooth_jac_and_residual_parity(hyg, meshes, name, L):
    gm, om = meshes[name]
    n = gm.ndofs(L)
    x0, b0, d0 = rnd(n, 3), rnd(n, 4), rnd(n, 5)
    op = hyg.P1ConstantOperator(gm, L, L)
    x, b, d = (hyg.Function(gm, L, L) for _ in range(3))
    x.upload(L, x0)
    b.upload(L, b0)
    d.upload(L, d0)
    op.smooth_jac(d, b, x, 2.0 / 3.0, L, hyg.INNER)
    got = d.download(L)
    oop = ol.Operator(om, L, L)
    ox, ob, od = (ol.Function(om, L, L) for _ in range(3))
    ox.set(L, x0)
    ob.set(L, b0)
    od.set(L, d0)
    oop.smooth_jac(od, ob, ox, 2.0 / 3.0, L, ol.INNER)
    assert rel_inf(got, od.get(L)) <= TOL64
    # residual r = b - A x on inner DoFs, untouched elsewhere
    r = hyg.Function(gm, L, L)
    r.upload(L, d0)
    op.residual(r, b, x, L, hyg.INNER)
    got = r.download(L)
    oo = ol.Function(om, L, L)
    oo.set(L, d0)
    oop.apply(ox, oo, L, ol.INNER, False)
    _, bnd = om.coords(L)
    ref = np.where(bnd, d0, b0 - oo.get(L))
    assert rel_inf(got, ref) <= TOL64
    # inverse diagonal
    dinv = hyg.Function(gm, L, L)
    op.inverse_diagonal(dinv, L)
    assert rel_inf(dinv.download(L), oop.inv_diag(L)) <= 1e-14


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("Lf", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("flag", ["inner", "all", "dirichlet"])
def test_transfer_parity(hyg, meshes, name, Lf, flag):
    """With a flag that excludes Inner the cell-interior points are still transferred (the reference's cell
    loops take no flag) and only the macro-face/-edge/-vertex points follow the flag."""
    gm, om = meshes[name]
    Lc = Lf - 1
    fl = {"inner": hyg.INNER, "all": hyg.ALL, "dirichlet": hyg.DIRICHLET}[flag]
    vc, vf = rnd(gm.ndofs(Lc), 21), rnd(gm.ndofs(Lf), 22)
    for op_name in ("restrict", "prolong_replace", "prolong_add"):
        g = hyg.Function(gm, Lc, Lf)
        o = ol.Function(om, Lc, Lf)
        g.upload(Lc, vc)
        g.upload(Lf, vf)
        o.set(Lc, vc)
        o.set(Lf, vf)
        if op_name == "restrict":
            g.restrict(Lf, fl)
            ol.lib().or_restrict(o.h, Lf, fl)
            got, ref = g.download(Lc), o.get(Lc)
        else:
            add = op_name == "prolong_add"
            g.prolongate(Lc, fl, hyg.ADD if add else hyg.REPLACE)
            ol.lib().or_prolongate(o.h, Lc, fl, int(add))
            got, ref = g.download(Lf), o.get(Lf)
        assert rel_inf(got, ref) <= TOL64, (op_name, name, Lf, flag)


@pytest.mark.parametrize("flag", ["inner", "all"])
def test_vector_ops_parity(hyg, meshes, flag):
    gm, om = meshes["cube_24el.msh"]
    L = 4
    n = gm.ndofs(L)
    fl = hyg.INNER if flag == "inner" else hyg.ALL
    a0, b0, c0 = rnd(n, 1), rnd(n, 2), rnd(n, 3)
    _, bnd = gm.coords(L)
    sel = ~bnd if flag == "inner" else np.ones(n, bool)
    a, b, c, d = (hyg.Function(gm, L, L) for _ in range(4))
    a.upload(L, a0)
    b.upload(L, b0)
    c.upload(L, c0)
    d.upload(L, c0)
    d.assign([2.0, -3.0, 0.5], [a, b, c], L, fl)
    ref = np.where(sel, 2 * a0 - 3 * b0 + 0.5 * c0, c0)
    assert rel_inf(d.download(L), ref) <= 1e-15
    d.add([1.5], [a], L, fl)
    ref = np.where(sel, ref + 1.5 * a0, ref)
    assert rel_inf(d.download(L), ref) <= 1e-15
    d.mult(a, b, L, fl)
    ref = np.where(sel, a0 * b0, ref)
    assert rel_inf(d.download(L), ref) <= 1e-15
    d.interpolate(7.0, L, fl)
    ref = np.where(sel, 7.0, ref)
    assert rel_inf(d.download(L), ref) <= 1e-15
    dot = a.dot(b, L, fl)
    refdot = float(np.sum((a0 * b0)[sel]))
    assert abs(dot - refdot) <= 1e-12 * np.sum(np.abs(a0 * b0))
    # value-type conversion round trip
    f32 = hyg.Function(gm, L, L, hyg.F32)
    f32.convert_from(a, L)
    assert np.array_equal(f32.download(L), a0.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("name", ["cube_kuhn_6el.msh", "cube_24el.msh"])
@pytest.mark.parametrize("smoother", ["jacobi", "cheb"])
def test_vcycle_residual_history_parity(hyg, meshes, name, smoother):
    gm, om = meshes[name]
    minL, maxL = 0, 5 if name == "cube_kuhn_6el.msh" else 4
    n = gm.ndofs(maxL)
    _, bnd = gm.coords(maxL)
    b0 = rnd(n, 123)
    b0[bnd] = 0.0
    kw = dict(pre=2, post=2, smoother="jacobi") if smoother == "jacobi" else \
        dict(pre=3, post=3, smoother="cheb", cheb_order=2, cheb_bounds=(0.3 * 1.9, 1.2 * 1.9))
    cycles = 6
    op = hyg.P1ConstantOperator(gm, minL, maxL)
    x, b = hyg.Function(gm, minL, maxL), hyg.Function(gm, minL, maxL)
    b.upload(maxL, b0)
    g = hyg.GMG(op, minL, maxL, **kw)
    got = g.solve(x, b, maxL, cycles)
    oop = ol.Operator(om, minL, maxL)
    ox, ob = ol.Function(om, minL, maxL), ol.Function(om, minL, maxL)
    ob.set(maxL, b0)
    ref = ol.gmg(oop, ox, ob, minL, maxL, cycles, **kw)
    assert np.all(np.abs(got - ref) / ref <= 1e-10), (got, ref)
    assert rel_inf(x.download(maxL), ox.get(maxL)) <= 1e-10


@pytest.mark.parametrize("L", [6, 7, 8])
def test_apply_parity_large(hyg, meshes, L):
    """Full-size apply on the 24-tet cube against the oracle, plus A*linear = 0 on inner DoFs."""
    gm, om = meshes["cube_24el.msh"]
    n = gm.ndofs(L)
    src = rnd(n, 1000 + L)
    op = hyg.P1ConstantOperator(gm, L, L)
    s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    s.upload(L, src)
    op.apply(s, d, L, hyg.INNER)
    got = d.download(L)
    oop = ol.Operator(om, L, L)
    os_, od = ol.Function(om, L, L), ol.Function(om, L, L)
    os_.set(L, src)
    oop.apply(os_, od, L, ol.INNER, False)
    assert rel_inf(got, od.get(L)) <= TOL64
    del os_, od, oop
    # the discrete Laplacian annihilates linear functions on inner DoFs (P1LaplaceOperator3DTest)
    xyz, bnd = gm.coords(L)
    lin = 0.3 + xyz @ np.array([1.0, -2.0, 0.5])
    s.upload(L, lin)
    op.apply(s, d, L, hyg.INNER)
    r = d.download(L)[~bnd]
    assert np.abs(r).max() <= 1e-13 * np.abs(lin).max() * (1 << L)


def test_apply_at_level_9(hyg, meshes):
    """The largest level the library takes (543M DoFs on the 24-tet cube, 4.3 GB per FP64 function), checked
    through size-independent properties of the assembled Laplacian: A 1 = 0 on inner DoFs, zero column sums
    (sum_i (A x)_i = 0 over all DoFs for the operator without boundary conditions), and symmetry
    (A x, z) = (x, A z); the Jacobi sweep with b = A x reproduces x on inner DoFs."""
    gm, _ = meshes["cube_24el.msh"]
    L = 9
    n = gm.ndofs(L)
    op = hyg.P1ConstantOperator(gm, L, L)
    x, y, z, w = (hyg.Function(gm, L, L) for _ in range(4))
    x.interpolate(1.0, L)
    op.apply(x, y, L, hyg.INNER)
    assert y.dot(y, L, hyg.INNER) <= 1e-24 * n
    rng = np.random.default_rng(909)
    x.upload(L, rng.uniform(-1.0, 1.0, n))
    z.upload(L, rng.uniform(-1.0, 1.0, n))
    op.apply(x, y, L, hyg.ALL)
    op.apply(z, w, L, hyg.ALL)
    ny = np.sqrt(y.dot(y, L))
    ones = hyg.Function(gm, L, L)
    ones.interpolate(1.0, L)
    assert abs(y.dot(ones, L)) <= 1e-12 * ny * np.sqrt(n)
    assert abs(y.dot(z, L) - x.dot(w, L)) <= 1e-12 * ny * np.sqrt(n)
    # one Jacobi sweep with b = A x started from x leaves x unchanged on inner DoFs (relative 1e-12)
    op.apply(x, y, L, hyg.INNER)
    w.assign([1.0], [x], L)
    op.smooth_jac(w, y, x, 0.8, L, hyg.INNER)
    w.add([-1.0], [x], L)
    assert np.sqrt(w.dot(w, L, hyg.INNER)) <= 1e-12 * np.sqrt(x.dot(x, L, hyg.INNER))


@pytest.mark.parametrize("cfg", range(7))
@pytest.mark.parametrize("L", [3, 5, 6])
def test_every_cell_kernel_configuration(hyg, meshes, cfg, L):
    """Apply, residual and Jacobi sweep under each cell-kernel configuration (rows x points per lane),
    FP64 against the oracle and FP32 against FP64 (relative 1e-5)."""
    gm, om = meshes["cube_24el.msh"]
    n = gm.ndofs(L)
    src, rhs = rnd(n, 300 + L), rnd(n, 400 + L)
    op = hyg.P1ConstantOperator(gm, L, L)
    for mode in range(6):
        op.set_config(L, hyg.F64, mode, cfg)
        op.set_config(L, hyg.F32, mode, cfg)
    oop = ol.Operator(om, L, L)
    os_, ob, od = ol.Function(om, L, L), ol.Function(om, L, L), ol.Function(om, L, L)
    os_.set(L, src)
    ob.set(L, rhs)
    for flag in (hyg.INNER, hyg.ALL):
        s, d = hyg.Function(gm, L, L), hyg.Function(gm, L, L)
        s.upload(L, src)
        op.apply(s, d, L, flag)
        od.set(L, np.zeros(n))
        oop.apply(os_, od, L, flag, False)
        assert rel_inf(d.download(L), od.get(L)) <= TOL64, (cfg, L, flag)
        s32, d32 = hyg.Function(gm, L, L, hyg.F32), hyg.Function(gm, L, L, hyg.F32)
        s32.upload(L, src)
        op.apply(s32, d32, L, flag)
        assert rel_inf(d32.download(L), od.get(L)) <= 1e-5, (cfg, L, flag, "fp32")
    # Jacobi sweep and residual on inner DoFs
    x, b, r = hyg.Function(gm, L, L), hyg.Function(gm, L, L), hyg.Function(gm, L, L)
    x.upload(L, src)
    b.upload(L, rhs)
    op.residual(r, b, x, L, hyg.INNER)
    od.set(L, np.zeros(n))
    oop.apply(os_, od, L, ol.INNER, False)
    _, bnd = om.coords(L)
    assert rel_inf(r.download(L), np.where(bnd, 0.0, rhs - od.get(L))) <= TOL64
    d = hyg.Function(gm, L, L)
    op.smooth_jac(d, b, x, 0.6, L, hyg.INNER)
    od.set(L, np.zeros(n))
    oop.smooth_jac(od, ob, os_, 0.6, L, ol.INNER)
    assert rel_inf(d.download(L), od.get(L)) <= TOL64
