"""CPU checks of the C ABI library (no device work): it loads, exports every entry point that
include/hyteg_amd.h declares, reports errors through status codes and hyg_last_error, and its host
side computations agree with the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol


def test_library_exports_every_declared_entry_point(hyg):
    lib = hyg.load()
    names = hyg.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding declares a signature for each of them
    lib2 = C.CDLL(hyg.LIB_PATH)
    table = hyg._sig(lib2)
    assert set(names) <= set(table), sorted(set(names) - set(table))


def test_calls_before_init_fail_with_a_state_error(hyg):
    lib = hyg.load()
    out = C.c_void_p()
    rc = lib.hyg_mesh_create_from_msh(b"meshes/cube_24el.msh", C.byref(out))
    assert rc == hyg.ERR_STATE
    assert b"hyg_init" in lib.hyg_last_error()
    assert not out.value


def test_null_handles_are_rejected(hyg):
    lib = hyg.load()
    assert lib.hyg_apply(None, None, None, 3, hyg.INNER, hyg.REPLACE, None) != hyg.OK
    assert len(lib.hyg_last_error()) > 0


def test_init_without_a_device_reports_an_error(hyg):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    rc = hyg.load().hyg_init(0)
    assert rc in (hyg.ERR_HIP, hyg.ERR_ARG)
    assert len(hyg.load().hyg_last_error()) > 0


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_chebyshev_coefficients_match_the_oracle(hyg, order):
    lo, hi = 0.3 * 1.96, 1.2 * 1.96
    a, b = hyg.cheb_coefficients(order, lo, hi), ol.cheb_coefficients(order, lo, hi)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()  # same closed form, different evaluation order


def test_chebyshev_order_out_of_range_is_an_argument_error(hyg):
    out = np.zeros(8)
    assert hyg.load().hyg_cheb_coefficients(9, 0.5, 2.0, out) == hyg.ERR_ARG


def test_fourier_field_matches_the_oracle_helper(hyg):
    for seed, m, k in ((2024, 32, 8), (3, 8, 3)):
        a, b = hyg.fourier_field(seed, modes=m, kmax=k), ol.random_field(seed, modes=m, kmax=k)
        assert a.shape == (m, 5) and np.array_equal(a, b)
        assert np.all(np.abs(a[:, :3]).max(axis=1) >= 1) and np.all(np.abs(a[:, :3]) <= k)
