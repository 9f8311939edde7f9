"""hyg_dot_noalloc (the dot product that allocates nothing) against hyg_dot, through the C ABI.

Both run the same kernels in the same order, so the results agree bit for bit; two calls in a row, the second
on a smaller input, show that the fixed scratch keeps no state between calls. A level that needs more partial
sums than the scratch holds is rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    return C.CDLL("libamdhip64.so")


def _device_double(hip):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), C.c_size_t(8)) == 0
    return p


def _read(hip, p):
    v = C.c_double()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(C.byref(v), p, C.c_size_t(8), 2) == 0  # device to host
    return v.value


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_dot_equals_dot_on_two_calls(hyg, dtype):
    hyg.init()
    lib = hyg.load()
    hip = _hip()
    dt = hyg.F64 if dtype == "f64" else hyg.F32
    out = _device_double(hip)
    try:
        cases = [("cube_24el.msh", 6, hyg.ALL), ("cube_kuhn_6el.msh", 3, hyg.INNER)]  # second call: smaller input
        for k, (mesh, L, flag) in enumerate(cases):
            print(f"no-alloc dot: {mesh} L{L} {dtype} flag {flag}", flush=True)
            m = hyg.Mesh(mesh)
            n = m.ndofs(L)
            rng = np.random.default_rng(40 + k)
            a, b = hyg.Function(m, L, L, dt), hyg.Function(m, L, L, dt)
            a.upload(L, rng.uniform(-1.0, 1.0, n))
            b.upload(L, rng.uniform(-1.0, 1.0, n))
            ref = a.dot(b, L, flag)
            rc = lib.hyg_dot_noalloc(a.h, b.h, L, flag, out, None)
            assert rc == hyg.OK, lib.hyg_last_error()
            got = _read(hip, out)
            assert got == ref, (mesh, L, got, ref)
    finally:
        hip.hipFree(out)


def test_noalloc_dot_rejects_levels_beyond_the_scratch(hyg):
    hyg.init()
    lib = hyg.load()
    hip = _hip()
    out = _device_double(hip)
    try:
        # cube24 level 9: 24 cells x ~20k plane tiles of partial sums, beyond the 262144 of the scratch
        m = hyg.Mesh("cube_24el.msh")
        L = 9
        a = hyg.Function(m, L, L, hyg.F32)
        rc = lib.hyg_dot_noalloc(a.h, a.h, L, hyg.ALL, out, None)
        assert rc == hyg.ERR_ARG
        assert b"fixed scratch" in lib.hyg_last_error()
        assert lib.hyg_dot_noalloc(a.h, a.h, L, hyg.ALL, None, None) == hyg.ERR_ARG
    finally:
        hip.hipFree(out)
