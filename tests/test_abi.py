"""The C ABI library loads and exports every function include/mfcfd.h declares, and
the Python binding binds exactly those (no GPU calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mfcfd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_api():
    names = declared_functions()
    for required in ("mf_create", "mf_destroy", "mf_run", "mf_cal_flux_residual",
                     "mf_state_update", "mf_eval_q_derivatives", "mf_func_delta",
                     "mf_compute_cl_cd_cm", "mf_get_prim", "mf_set_prim"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from mfcfd_amd import _lib
    lib = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    from mfcfd_amd import _lib
    bound = sorted(n for n, _, _ in _lib.SIGNATURES)
    assert bound == declared_functions()


def test_struct_layouts():
    """mf_geometry / mf_params sizes match the C declarations (ctypes mirror)."""
    from mfcfd_amd import _lib
    assert C.sizeof(_lib.MfGeometry) == 8 + 11 * 8
    assert C.sizeof(_lib.MfParams) == 7 * 8 + 10 * 4


def test_default_params_and_error_path():
    from mfcfd_amd import _lib
    from mfcfd_amd.solver import make_params
    p = make_params()
    assert p.rks == 4 and p.euler == 1.0 and p.vl_const == 150.0 and p.power == 0.0
    p = make_params(tscheme="first", precision="mixed", cfl=0.3)
    assert p.rks == 1 and p.euler == 2.0 and p.precision == _lib.MF_MIXED and p.cfl == 0.3
    with pytest.raises(KeyError):
        make_params(no_such_option=1)
    # invalid geometry is rejected on the host before any device call
    g = _lib.MfGeometry(n_points=0, max_nbhs=24)
    h = C.c_void_p()
    rc = _lib.lib().mf_create(C.byref(g), C.byref(make_params()), C.byref(h))
    assert rc == -2 and b"n_points" in _lib.lib().mf_last_error()


def test_invalid_cloud_rejected(small_cloud):
    """Out-of-range neighbour indices and bad flags fail in mf_create's validation."""
    from mfcfd_amd import _lib
    from mfcfd_amd.solver import LskumSolver
    import copy
    bad = copy.deepcopy(small_cloud)
    bad.conn[5, 0] = bad.n + 3
    with pytest.raises(_lib.MfError, match="out of range"):
        LskumSolver(bad)
    bad = copy.deepcopy(small_cloud)
    bad.flag_1[7] = 5
    with pytest.raises(_lib.MfError, match="flag_1"):
        LskumSolver(bad)
    bad = copy.deepcopy(small_cloud)
    bad.flag_2[0] = 3
    with pytest.raises(_lib.MfError, match="shapes"):
        LskumSolver(bad)
