"""CPU checks of the hand-written reverse functions (mfcfd_amd/csrc/lskum_adjoint.h).

Each reverse function is compiled for the host (tests/native/adjcheck.hip) next to
its forward, and checked by the dot-product identity w . (J v) = (J^T w) . v with
J v from central finite differences of the forward, for random states, directions v
and seeds w.  Branches (limiter thresholds, phi capped at 1) are kept away from the
finite-difference stencil.
"""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libadjcheck.so")


@pytest.fixture(scope="module")
def ac():
    if not os.path.exists(LIB):
        import importlib.util
        spec = importlib.util.spec_from_file_location("adjbuild", os.path.join(HERE, "native", "build.py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m.build()
    lib = C.CDLL(LIB)
    lib.ac_wave_fwd.restype = C.c_double
    lib.ac_wave_rev.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    lib.ac_update_fwd.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.ac_update_rev.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def _ptr(a):
    return a.ctypes.data


def dot_test(fwd, rev, x, nout, rng, h=1e-6, trials=4, rtol=2e-6, vmask=None):
    x = np.asarray(x, dtype=np.float64)
    for _ in range(trials):
        v = rng.standard_normal(x.shape)
        if vmask is not None:
            v = v * vmask
        w = rng.standard_normal(nout)
        fp, fm = fwd(x + h * v), fwd(x - h * v)
        jv = (fp - fm) / (2 * h)
        lhs = float(w @ jv)
        rhs = float(rev(x, w) @ v)
        scale = max(abs(lhs), abs(rhs), 1e-12, np.abs(jv).max() * np.abs(w).max() * 1e-3)
        assert abs(lhs - rhs) <= rtol * scale + 1e-9, (lhs, rhs)


def _qt_state(rng):
    rho, u1, u2, p = rng.uniform(0.6, 1.4), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.4, 1.0)
    beta = 0.5 * rho / p
    q = np.array([np.log(rho) + 2.5 * np.log(beta) - beta * (u1 * u1 + u2 * u2), 2 * beta * u1, 2 * beta * u2, -2 * beta])
    th = rng.uniform(0, 2 * np.pi)
    return np.concatenate([q, [np.cos(th), np.sin(th)]])


@pytest.mark.parametrize("kind", range(8))
def test_flux_and_kin_reverse(ac, kind):
    rng = np.random.default_rng(100 + kind)

    def fwd(x):
        G = np.zeros(4)
        ac.ac_flux_fwd(kind, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(G)))
        return G

    def rev(x, w):
        out = np.zeros(6)
        w = np.ascontiguousarray(w)
        ac.ac_flux_rev(kind, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(w)), C.c_void_p(_ptr(out)))
        return out

    for _ in range(5):
        x = _qt_state(rng)
        dot_test(fwd, rev, x, 4, rng)


def _recon_state(rng, active):
    q = rng.uniform(-1, 1, 4)
    dx = rng.uniform(-0.5, 0.5, 4)
    dy = rng.uniform(-0.5, 0.5, 4)
    spread = rng.uniform(0.002, 0.02, 4)
    qmax, qmin = q + spread, q - spread
    eps = 1e-3
    # active: |dneg| well above the 1e-5 threshold; inactive: below it (phi = 1)
    delx, dely = rng.uniform(-0.02, 0.02, 2) if active else rng.uniform(-1e-5, 1e-5, 2)
    return np.concatenate([q, dx, dy, qmax, qmin, [eps, delx, dely]])


def _far_from_limiter_branches(x, margin=1e-3):
    """True when no component sits near a branch of the limiter: the |dneg| threshold,
    dneg = 0, or the cap phi = min(t, 1)."""
    q, dx, dy, qmax, qmin = x[0:4], x[4:8], x[8:12], x[12:16], x[16:20]
    eps, delx, dely = x[20], x[21], x[22]
    d = delx * dx + dely * dy
    dn = -0.5 * d
    if np.any(np.abs(dn) < 1e-4):
        return False
    dp = np.where(dn > 0, qmax - q, qmin - q)
    num = (dp * dp + eps * eps) * dn + 2 * dn * dn * dp
    den = (dp * dp + 2 * dn * dn + dn * dp + eps * eps) * dn
    return bool(np.all(np.abs(num / den - 1.0) > margin))


@pytest.mark.parametrize("active", [True, False])
def test_reconstruct_reverse(ac, active):
    rng = np.random.default_rng(7 if active else 8)

    def fwd(x):
        out = np.zeros(4)
        ac.ac_recon_fwd(C.c_void_p(_ptr(x)), C.c_void_p(_ptr(out)))
        return out

    def rev(x, w):
        out = np.zeros(23)
        w = np.ascontiguousarray(w)
        ac.ac_recon_rev(C.c_void_p(_ptr(x)), C.c_void_p(_ptr(w)), C.c_void_p(_ptr(out)))
        return out

    done = 0
    for _ in range(500):
        if done == 6:
            break
        x = _recon_state(rng, active)
        if active and not _far_from_limiter_branches(x):
            continue
        if not active and np.any(np.abs(0.5 * (x[21] * x[4:8] + x[22] * x[8:12])) > 6e-6):
            continue
        done += 1
        v_mask = np.ones(23)
        v_mask[20] = 0.0  # epsilon is a constant of the point, not differentiated
        dot_test(fwd, rev, x, 4, rng, h=1e-7 if active else 1e-8, vmask=v_mask)
    assert done == 6


@pytest.mark.parametrize("cls,rk", [(0, 1), (0, 3), (1, 2), (1, 3), (2, 1), (2, 3)])
def test_update_reverse(ac, cls, rk):
    rng = np.random.default_rng(10 * cls + rk)
    fs = np.array([0.6, 0.05])

    def fwd(x):
        out = np.zeros(4)
        ac.ac_update_fwd(cls, rk, 1.0, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(fs)), C.c_void_p(_ptr(out)))
        return out

    def rev(x, w):
        out = np.zeros(14)
        w = np.ascontiguousarray(w)
        ac.ac_update_rev(cls, rk, 1.0, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(fs)), C.c_void_p(_ptr(w)),
                         C.c_void_p(_ptr(out)))
        return out

    for _ in range(4):
        p = np.array([rng.uniform(0.8, 1.2), rng.uniform(-0.7, 0.7), rng.uniform(-0.3, 0.3), rng.uniform(0.6, 0.8)])
        po = p * rng.uniform(0.97, 1.03, 4)
        fr = rng.uniform(-0.01, 0.01, 4)
        th = rng.uniform(0, 2 * np.pi)
        x = np.concatenate([p, po, fr, [np.cos(th), np.sin(th)]])
        dot_test(fwd, rev, x, 4, rng)


def test_q_of_prim_and_wave_speed_reverse(ac):
    rng = np.random.default_rng(3)

    def qf(x):
        out = np.zeros(4)
        ac.ac_qprim_fwd(C.c_void_p(_ptr(x)), C.c_void_p(_ptr(out)))
        return out

    def qr(x, w):
        out = np.zeros(4)
        w = np.ascontiguousarray(w)
        ac.ac_qprim_rev(C.c_void_p(_ptr(x)), C.c_void_p(_ptr(w)), C.c_void_p(_ptr(out)))
        return out

    def wf(x):
        return np.array([ac.ac_wave_fwd(C.c_void_p(_ptr(x)))])

    def wr(x, w):
        out = np.zeros(4)
        ac.ac_wave_rev(C.c_void_p(_ptr(x)), float(w[0]), C.c_void_p(_ptr(out)))
        return out

    for _ in range(5):
        p = np.array([rng.uniform(0.8, 1.2), rng.uniform(-0.7, 0.7), rng.uniform(-0.3, 0.3), rng.uniform(0.6, 0.8)])
        dot_test(qf, qr, p, 4, rng)
        dot_test(wf, wr, p, 1, rng)
