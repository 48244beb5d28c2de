"""The hand-written reverse functions against the reference's Tapenade adjoint,
routine by routine (CPU).

tests/golden/tap_routines.npz holds outputs of the reference's own Tapenade-generated
reverse routines (src_mpi_adjoint/sensitivity_files, compiled unmodified into
oracle/_tap by oracle/Makefile; tests/golden/make_tap_golden.py).

* Point routines, directly: the host build of mfcfd_amd/csrc/lskum_adjoint.h
  (tests/native/adjcheck.hip) against FLUX_*_B composed with QTILDE_TO_PRIMITIVE_B
  (the eight kinetic fluxes, split_fluxes_b.F90 and quadrant_fluxes_b.F90),
  VENKAT_LIMITER_B (limiters_b.F90:43-164) and EVAL_Q_VARIABLES_B
  (q_variables_b.F90:13-50), to 1e-12 relative (the limiter's derivative cancels terms
  of similar size: where the two evaluations differ by more, both are held to a
  50-digit evaluation of it).
* Cloud routines, by the dot-product identity w . (J v) = (J^T w) . v with J v from
  the oracle's dual-number tangent of the same routine (the tangent the GPU adjoint is
  checked against over whole iterations): CAL_FLUX_RESIDUAL_B
  (flux_residual_b.F90:23-120), EVAL_Q_DERIVATIVES_B (q_variables_b.F90:80-253), one
  inner iteration (EVAL_UPDATE_INNERLOOP_B, EVAL_Q_INNER_LOOP_B, :313-767) and
  COMPUTE_NORMALS_B (point_normals_b.F90:15-162), to 1e-12 relative.
* Where oracle/_tap is built, the fixture is recomputed and must match bit for bit.

The whole-iteration pointb%x/y stays unpinned against Tapenade itself: state_update_b,
fpi_solver_b and compute_force_coeffs_b include PETSc headers and cannot be built here.
"""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "tap_routines.npz")
LIB = os.path.join(HERE, "native", "libadjcheck.so")


@pytest.fixture(scope="module")
def tap():
    return np.load(GOLD, allow_pickle=False)


@pytest.fixture(scope="module")
def ac():
    if not os.path.exists(LIB):
        import importlib.util
        spec = importlib.util.spec_from_file_location("adjbuild", os.path.join(HERE, "native", "build.py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m.build()
    lib = C.CDLL(LIB)
    lib.ac_flux_rev.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.ac_venkat_rev.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.ac_qprim_rev.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def _p(a):
    return a.ctypes.data


def _close(a, b, rtol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rtol * max(np.abs(b).max(), 1e-300)


def test_fluxes_against_tapenade(ac, tap):
    fin, gb, fout = tap["flux_in"], tap["flux_gb"], tap["flux_out"]
    for i in range(fin.shape[0]):
        x = np.ascontiguousarray(fin[i])
        g = np.ascontiguousarray(gb[i])
        for k in range(8):
            out = np.zeros(6)
            ac.ac_flux_rev(k, _p(x), _p(g), _p(out))
            assert _close(out, fout[i, k]), (i, k, out, fout[i, k])


def test_limiter_against_tapenade(ac, tap):
    from tapenade_state import golden_state
    st, vl, min_dist = golden_state(tap)
    n_lim = n_cancel = 0
    for i in range(tap["venkat_k"].shape[0]):
        k = int(tap["venkat_k"][i])
        e = vl * min_dist[k]
        eps = (e * e) * e  # amdflang's lowering of epsi**3.0d0 (limiters_b.F90:84-85)
        inp = np.concatenate([tap["venkat_qt"][i], st["q"][k], st["qm"][k, :, 0], st["qm"][k, :, 1], [eps]])
        out = np.zeros(16)
        ac.ac_venkat_rev(_p(inp), _p(np.ascontiguousarray(tap["venkat_phib"][i])), _p(out))
        ref = tap["venkat_out"][i]
        for j in range(16):
            if abs(out[j] - ref[j]) > 1e-12 * max(abs(ref[j]), 1e-300):
                # the ratio's derivative cancels terms of similar size, so two FP64
                # evaluations in different orders differ by more than their rounding:
                # there both are held to the derivative evaluated with 50 digits, the
                # host reverse within 10 times Tapenade's own error of it or 1e-11
                exact = _limiter_derivative(inp, tap["venkat_phib"][i], j)
                err_t, err_h = abs(ref[j] - exact), abs(out[j] - exact)
                assert err_t <= 1e-6 * abs(exact), (i, j)
                assert err_h <= max(10.0 * err_t, 1e-11 * abs(exact)), (i, j, err_h, err_t)
                n_cancel += 1
        n_lim += np.abs(out[:4]).max() > 0
    assert n_lim >= 16  # the limited branch is exercised, not only phi = 1
    print(f"limiter: {n_lim} limited cases, {n_cancel} components held to the 50-digit derivative")


def _limiter_derivative(inp, phib, j):
    """d(sum_r phib_r phi_r)/d(input j) of venkat_limiter with 50-digit arithmetic; j
    indexes (q-tilde 4, q 4, qmax 4, qmin 4) as ac_venkat_rev's output."""
    import mpmath as mp
    mp.mp.dps = 50
    r = j % 4
    qt, q, qmax, qmin, eps = (mp.mpf(float(v)) for v in (inp[r], inp[4 + r], inp[8 + r], inp[12 + r], inp[16]))

    def phi(qt_, q_, qmax_, qmin_):
        dn = qt_ - q_
        dp = (qmax_ if dn > 0 else qmin_) - q_
        num = (dp * dp + eps * eps) * dn + 2 * dn * dn * dp
        den = (dp * dp + 2 * dn * dn + dn * dp + eps * eps) * dn
        return min(mp.mpf(1), num / den)
    args = [qt, q, qmax, qmin]
    which = j // 4
    d = mp.diff(lambda z: phi(*[z if a == which else args[a] for a in range(4)]), args[which])
    return float(mp.mpf(float(phib[r])) * d)


def test_q_variables_against_tapenade(ac, tap):
    prim = tap["st_prim"]
    qb = tap["qvar_qb"].reshape(-1, 4)
    ref = tap["qvar_primb"].reshape(-1, 4)
    for k in range(prim.shape[0]):
        pb = np.zeros(4)
        ac.ac_qprim_rev(_p(np.ascontiguousarray(prim[k])), _p(np.ascontiguousarray(qb[k])), _p(pb))
        assert _close(pb, ref[k]), k


def _dot_check(d, setup, outputs, adjoints, seeds, rng, routines):
    """Sets each field of `setup` with a random tangent, calls the routines, and checks
    sum(seed . output tangent) against sum(adjoint . input tangent)."""
    V = {}
    for name, val in setup.items():
        V[name] = rng.standard_normal(np.asarray(val).size) if name in adjoints else None
        d.set(name, val, V[name])
    for r in routines:
        d.call(r)
    lhs = sum(float(seeds[o] @ d.get(o)[1]) for o in outputs)
    rhs = sum(float(adjoints[a] @ V[a]) for a in adjoints)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)
    return lhs


@pytest.fixture(scope="module")
def dual(tap):
    from tapenade_state import golden_dual
    return golden_dual()


def test_flux_residual_against_tapenade(dual, tap):
    st = {k: tap["st_" + k] for k in ("x", "y", "nx", "ny", "q", "dq", "qm", "delta")}
    adj = {k: tap["cfr_" + k] for k in st}
    lhs = _dot_check(dual, st, ["flux_res"], adj, {"flux_res": tap["cfr_w"]}, np.random.default_rng(1),
                     ["cal_flux_residual"])
    assert abs(lhs) > 1.0


def test_q_derivatives_against_tapenade(dual, tap):
    st = {k: tap["st_" + k] for k in ("x", "y", "q")}
    adj = {k: tap["der_" + k] for k in st}
    _dot_check(dual, st, ["dq", "qm"], adj, {"dq": tap["der_wdq"], "qm": tap["der_wqm"]},
               np.random.default_rng(2), ["eval_q_derivatives"])


def test_inner_iteration_against_tapenade(dual, tap):
    st = {"x": tap["st_x"], "y": tap["st_y"], "q": tap["st_q"], "dq": tap["st_dq0"]}
    adj = {k: tap["inn_" + k] for k in st}
    _dot_check(dual, st, ["dq"], adj, {"dq": tap["inn_w"]}, np.random.default_rng(3),
               ["eval_q_inner_loop", "eval_update_innerloop"])


def test_normals_against_tapenade(dual, tap):
    st = {"x": tap["st_x"], "y": tap["st_y"]}
    adj = {"x": tap["nrm_x"], "y": tap["nrm_y"]}
    _dot_check(dual, st, ["nx", "ny"], adj, {"nx": tap["nrm_wx"], "ny": tap["nrm_wy"]},
               np.random.default_rng(4), ["compute_normals"])


def test_live_tapenade_matches_fixture(tap):
    from oracle.oracle import TapRoutines
    if not TapRoutines.available():
        pytest.skip("reference sources not built here (oracle/_tap)")
    import importlib.util
    spec = importlib.util.spec_from_file_location("mtg", os.path.join(HERE, "golden", "make_tap_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    live = m.compute()
    assert sorted(live) == sorted(tap.files)
    for k in tap.files:
        assert np.array_equal(live[k], tap[k]), k
