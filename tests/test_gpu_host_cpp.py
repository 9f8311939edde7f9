"""The C++ host classes (hyteg-vara_amd/host/hyteg_amd.hpp) driven by a plain C++ program, against the oracle: P1
apply and V-cycle residual history, and the P2 apply."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hyteg-vara_amd", "gmg_example")


def test_cpp_host_program_matches_oracle(tmp_path):
    assert os.path.exists(EXE), "gmg_example is not built"
    name, L, cycles = "cube_24el.msh", 4, 3
    om = ol.Mesh(name)
    n = om.ndofs(L)
    v = np.random.default_rng(31).uniform(-1.0, 1.0, n)
    inp = tmp_path / "v.bin"
    v.tofile(inp)
    out = subprocess.run([EXE, os.path.join(ROOT, "meshes", name), str(L), str(cycles), str(inp)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout.strip().splitlines()[-1])
    assert got["n"] == n
    _, bnd = om.coords(L)
    # apply norm over inner DoFs
    op = ol.Operator(om, L, L)
    s, d = ol.Function(om, L, L), ol.Function(om, L, L)
    s.set(L, v)
    op.apply(s, d, L, ol.INNER, False)
    ref = np.sqrt(np.sum(d.get(L)[~bnd] ** 2))
    assert abs(got["apply_norm"] - ref) <= 1e-12 * ref
    # Jacobi V(2,2) residual history
    b0 = v.copy()
    b0[bnd] = 0.0
    op2 = ol.Operator(om, 0, L)
    x, b = ol.Function(om, 0, L), ol.Function(om, 0, L)
    b.set(L, b0)
    res = ol.gmg(op2, x, b, 0, L, cycles, pre=2, post=2, smoother="jacobi")
    assert np.all(np.abs(np.array(got["residuals"]) - res) <= 1e-10 * res), (got["residuals"], res)
    # P2ConstantLaplaceOperator through the C++ classes: P2 level L - 1 on the same values
    op3 = ol.P2Operator(om, L - 1, L - 1)
    s3, d3 = ol.Function(om, L, L), ol.Function(om, L, L)
    s3.set(L, v)
    op3.apply(s3, d3, L, ol.INNER, False)
    ref3 = np.sqrt(np.sum(d3.get(L)[~bnd] ** 2))
    assert abs(got["p2_apply_norm"] - ref3) <= 1e-12 * ref3
