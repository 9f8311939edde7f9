"""Error paths of the C ABI (the reference aborts or throws; here every failure is a status code).

  * HYG_ERR_ARG: an operator or solver paired with functions of another mesh (the kernels would walk the
    operator's cells over arrays sized for the other mesh), checked before any launch.
  * HYG_ERR_NUMERIC, forced from the host so that no kernel faults: a degenerate (flat) macro-cell gives a zero
    stencil centre at operator creation; a NaN in the right-hand side or the start vector gives a non-finite
    residual norm or spectral radius (NaN is an ordinary value for the kernels).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLAT_TET = """$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
4
1 0 0 0
2 1 0 0
3 0 1 0
4 0.25 0.25 0
$EndNodes
$Elements
1
1 4 2 1 1 1 2 3 4
$EndElements
"""


@pytest.fixture(scope="module")
def two(hyg):
    hyg.init()
    return hyg.Mesh("cube_24el.msh"), hyg.Mesh("cube_kuhn_6el.msh")


def _code(hyg, fn, *args):
    with pytest.raises(hyg.HygError) as e:
        fn(*args)
    return e.value.code


def test_operator_and_functions_of_different_meshes(hyg, two, hip_stream):
    a, b = two
    L = 3
    op = hyg.P1ConstantOperator(a, 1, L)
    x, y, z = (hyg.Function(b, 1, L) for _ in range(3))
    assert _code(hyg, op.apply, x, y, L) == hyg.ERR_ARG
    assert _code(hyg, op.smooth_jac, x, y, z, 0.6, L) == hyg.ERR_ARG
    assert _code(hyg, op.smooth_sor, x, y, 1.0, L) == hyg.ERR_ARG
    assert _code(hyg, op.residual, x, y, z, L) == hyg.ERR_ARG
    assert _code(hyg, op.inverse_diagonal, x, L) == hyg.ERR_ARG
    assert _code(hyg, op.power_iteration, x, y, 5, L) == hyg.ERR_ARG
    g = hyg.GMG(op, 1, L, pre=1, post=1)
    assert _code(hyg, g.vcycle, x, y, L) == hyg.ERR_ARG
    assert _code(hyg, g.vcycle_graph, x, y, L, 4, hip_stream) == hyg.ERR_ARG
    # defect correction: FP64 operator and FP32 solver of different meshes
    op_b = hyg.P1ConstantOperator(b, 1, L)
    g32 = hyg.GMG(op, 1, L, hyg.F32)
    assert _code(hyg, hyg.ir_solve, op_b, g32, x, y, L) == hyg.ERR_ARG
    # the same calls on matching meshes succeed
    xa, ya, za = (hyg.Function(a, 1, L) for _ in range(3))
    op.smooth_jac(xa, ya, za, 0.6, L)
    op.residual(xa, ya, za, L)


def test_degenerate_cell_gives_a_numeric_error(hyg, tmp_path):
    hyg.init()
    path = os.path.join(str(tmp_path), "flat_tet.msh")
    with open(path, "w") as fh:
        fh.write(FLAT_TET)
    m = hyg.Mesh(path)
    assert _code(hyg, hyg.P1ConstantOperator, m, 0, 3) == hyg.ERR_NUMERIC


def test_nan_inputs_give_a_numeric_error(hyg, two):
    a, _ = two
    minL, L = 0, 4
    n = a.ndofs(L)
    _, bnd = a.coords(L)
    b0 = np.random.default_rng(2).uniform(-1, 1, n)
    b0[bnd] = 0.0
    b0[np.flatnonzero(~bnd)[7]] = np.nan
    op = hyg.P1ConstantOperator(a, minL, L)
    x, b = hyg.Function(a, minL, L), hyg.Function(a, minL, L)
    b.upload(L, b0)
    g = hyg.GMG(op, minL, L, pre=1, post=1)
    assert _code(hyg, g.solve, x, b, L, 2) == hyg.ERR_NUMERIC
    v, t = hyg.Function(a, L, L), hyg.Function(a, L, L)
    v.upload(L, b0)
    assert _code(hyg, op.power_iteration, v, t, 5, L) == hyg.ERR_NUMERIC
