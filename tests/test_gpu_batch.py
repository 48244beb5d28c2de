"""Batched cases (mf_create_batch): several clouds with their own Mach and angle of
attack in one context must reproduce each case run on its own context."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

COMMON = dict(cfl=0.2, vl_const=5.0, inner_iterations=2)


@pytest.fixture(scope="module")
def cases(small_cloud):
    from mfcfd_amd.cloud import naca_ogrid
    b = naca_ogrid(40, 12, seed=9, thickness=0.10, first_spacing=0.01, far_radius=4.0)
    c = naca_ogrid(56, 10, seed=4, thickness=0.15, first_spacing=0.012, far_radius=4.0)
    return [(small_cloud, 0.63, 2.0), (b, 0.75, -1.0), (c, 0.5, 3.5)]


def test_batch_matches_single_cases(cases):
    from mfcfd_amd.solver import LskumBatch, LskumSolver
    iters = 20
    with LskumBatch(cases, **COMMON) as bt:
        res, rn, cd = bt.run(iters)
        prim = bt.prim
        cl_b, cd_b, cm_b, _ = bt.compute_cl_cd_cm()
        print(f"[batch] {bt.ncases} cases, {bt.n} points", flush=True)
        for k, (cloud, mach, aoa) in enumerate(cases):
            with LskumSolver(cloud, mach=mach, aoa=aoa, **COMMON) as s:
                r, rr, c = s.q_lskum(iters)
                assert rel_err(prim[bt.case_slice(k)], s.prim) <= 1e-14, k
                assert np.abs(res[:, k] - r).max() <= 1e-12, k
                assert np.abs(cd[:, k, 0] - c).max() <= 1e-13 * np.abs(c).max(), k
                cl, cdd, cm, _ = s.compute_cl_cd_cm()
                assert abs(cl_b[k, 0] - cl[0]) <= 1e-13 * abs(cl[0]) + 1e-16
                assert abs(cd_b[k, 0] - cdd[0]) <= 1e-13 * abs(cdd[0])


def test_batch_adjoint_is_per_case_gradient(cases):
    """J of a batch is the sum of the cases' Cd; each point's gradient is its own
    case's dCd/dx."""
    from mfcfd_amd.solver import LskumBatch, LskumSolver
    with LskumBatch(cases, **COMMON) as bt:
        J, dx, dy, _ = bt.adjoint(10, snaps=3)
        Js = 0.0
        for k, (cloud, mach, aoa) in enumerate(cases):
            with LskumSolver(cloud, mach=mach, aoa=aoa, **COMMON) as s:
                j, gx, gy, _ = s.adjoint(10, snaps=3)
            Js += j
            sl = bt.case_slice(k)
            g = np.concatenate([gx, gy])
            gb = np.concatenate([dx[sl], dy[sl]])
            assert np.abs(gb - g).max() <= 1e-10 * np.abs(g).max(), k
        assert abs(J - Js) <= 1e-13 * abs(Js)


def test_batch_mixed_precision_close(cases):
    from mfcfd_amd.solver import LskumBatch
    with LskumBatch(cases, **COMMON) as a, LskumBatch(cases, precision="mixed", **COMMON) as b:
        a.run(30)
        b.run(30)
        cla, cda, _, _ = a.compute_cl_cd_cm()
        clb, cdb, _, _ = b.compute_cl_cd_cm()
        assert np.abs(cla - clb).max() <= 1e-4 and np.abs(cda - cdb).max() <= 1e-4
