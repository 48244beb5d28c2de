"""GPU: the resident inner sweeps (every inner iteration of a stage in one launch of
the staged inner kernel's blocks, all resident at once, with a grid barrier between
sweeps; q_variables.F90:115-227) against the staged per-sweep kernel.

Both share the edge function, each lane's slot order and the solve, so trajectories are
compared bit for bit.  The barrier's give-up path is forced from the host with a
one-poll limit, which every kernel still exits from normally: the run stops, mf_run
reports the error, and the context falls back to the staged kernel with the same lanes.
"""
import numpy as np
import pytest

from mfcfd_amd import workloads as W

pytestmark = pytest.mark.gpu

STAGED, RESIDENT = 16, 32


@pytest.fixture(scope="module")
def c3_cloud():
    """The 320 x 125 NACA0012 cloud (40,000 points, 625 blocks of 64 points)."""
    return W.cloud_of(W.C3)


@pytest.fixture(scope="module")
def c2_cloud():
    """The 640 x 250 NACA0012 cloud (160,000 points, 2,500 blocks of 64 points)."""
    return W.cloud_of(W.C2)


def _mixed(cloud, params, **kw):
    from mfcfd_amd.solver import LskumSolver
    p = dict(params)
    p.update(kw)
    return LskumSolver(cloud, precision="mixed", **p)


def _pair(cloud, params, lanes, inner):
    """Two mixed contexts with the same LSQ and flux shapes: staged inner sweeps in the
    first, resident ones in the second (None if they do not fit this device)."""
    from mfcfd_amd._lib import MfError
    a = _mixed(cloud, params, inner_iterations=inner)
    b = _mixed(cloud, params, inner_iterations=inner)
    shapes = {"lpp_lsq": STAGED | 2, "lpp_flux": STAGED | 4}
    a.tuning = {**shapes, "lpp_inner": STAGED | lanes}
    try:
        b.tuning = {**shapes, "lpp_inner": RESIDENT | lanes}
    except MfError:
        a.close()
        b.close()
        return None
    assert b.tuning["lpp_inner"] == RESIDENT | lanes
    return a, b


@pytest.mark.parametrize("inner", [1, 2, 3])
@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_resident_sweeps_bit_identical(c3_cloud, lanes, inner):
    pair = _pair(c3_cloud, W.C3["params"], lanes, inner)
    if pair is None:
        pytest.skip(f"resident sweeps with {lanes} lanes do not fit every block of the cloud on this device")
    a, b = pair
    with a, b:
        print(f"[sweeps] 40,000 points, {lanes} lanes, {inner} inner iterations, 30 iterations", flush=True)
        ra, _, cda = a.q_lskum(30)
        rb, _, cdb = b.q_lskum(30)
        assert len(rb) == 30
        assert np.array_equal(a.prim, b.prim)
        assert np.array_equal(cda, cdb)
        assert np.array_equal(ra, rb)
        # the single-sweep routine on a resident context runs the staged sweep
        a.eval_q_derivatives()
        b.eval_q_derivatives()
        a.eval_q_inner_loop()
        b.eval_q_inner_loop()
        assert np.array_equal(a.field("dq"), b.field("dq"))


def test_resident_sweeps_full_cloud(c2_cloud):
    """The 160,000-point cloud: 2,500 blocks must all be resident.  At least one lane
    count fits on an MI355X; every one that fits matches the staged kernel."""
    fitted = []
    for lanes in (1, 2, 4):
        pair = _pair(c2_cloud, W.C2["params"], lanes, W.C2["params"]["inner_iterations"])
        if pair is None:
            continue
        fitted.append(lanes)
        a, b = pair
        with a, b:
            print(f"[sweeps] 160,000 points, {lanes} lanes, 3 inner iterations, 20 iterations", flush=True)
            a.q_lskum(20)
            b.q_lskum(20)
            assert np.array_equal(a.prim, b.prim)
    assert fitted, "no lane count of the resident sweeps fits the 160,000-point cloud"


def test_barrier_give_up_falls_back(c3_cloud):
    """A one-poll barrier limit: the first wait gives up, every kernel of the run
    returns, mf_run reports the error, and the context then runs the staged sweeps
    (same lanes) from a reloaded state exactly like a context that never used the
    resident ones."""
    from mfcfd_amd._lib import MfError
    pair = _pair(c3_cloud, W.C3["params"], 2, 2)
    if pair is None:
        pytest.skip("resident sweeps with 2 lanes do not fit every block of the cloud on this device")
    a, b = pair
    with a, b:
        start = a.prim.copy()
        b.tuning = {"barrier_polls": 1}
        assert b.tuning["barrier_polls"] == 1
        print("[sweeps] give-up: 40,000 points, 2 lanes, one-poll limit", flush=True)
        with pytest.raises(MfError, match="grid barrier"):
            b.q_lskum(5)
        assert b.tuning["lpp_inner"] == STAGED | 2
        b.prim = start
        a.q_lskum(5)
        b.q_lskum(5)
        assert np.array_equal(a.prim, b.prim)
        # back to the resident shape with the default limit: runs normally again
        b.tuning = {"lpp_inner": RESIDENT | 2, "barrier_polls": 0}
        a.q_lskum(5, first_it=6)
        b.q_lskum(5, first_it=6)
        assert np.array_equal(a.prim, b.prim)


def test_autotune_offers_resident(c3_cloud):
    """mf_autotune times the resident sweeps among the inner shapes (per sweep) and
    keeps a valid shape."""
    with _mixed(c3_cloud, W.C3["params"]) as s:
        t = s.autotune(reps=5)
        print(f"[sweeps] autotune: inner {t['lpp_inner']} at {t['us_inner']:.2f} us per sweep", flush=True)
        assert t["lpp_inner"] in (1, 2, 4, 8) + tuple(STAGED | v for v in (1, 2, 4)) + \
            tuple(RESIDENT | v for v in (1, 2, 4))
        assert t["us_inner"] > 0
