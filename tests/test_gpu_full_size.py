"""GPU parity on the five workloads at their full sizes and lengths (definitions in
mfcfd_amd/workloads.py; fpi_solver.F90:13-83 semantics throughout).

* 9,600 points, first order: FP64 against the oracle over all 1000 iterations.
* 160,000 points: FP64 against the oracle for 50 iterations; the mixed path against
  FP64 after all 5000 iterations (|dCl|, |dCd| <= 1e-4), under the autotuned shapes;
  staged walks against global walks (same bits).
* 2,560,000 points: FP64 against the oracle for 2 iterations; the mixed path against
  FP64 after the workload's 600 iterations (100 + 500).
* 64 batched cases (20k-120k points each): every case of the batch against its own
  single-case context; the smallest and the largest cloud against the oracle; the
  mixed batch against the FP64 batch after 1000 iterations, per case.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _gpu(cloud, params, **kw):
    from mfcfd_amd.solver import LskumSolver
    p = dict(params)
    p.update(kw)
    return LskumSolver(cloud, **p)


def _oracle(cloud, params):
    from mfcfd_amd.workloads import oracle_params
    from oracle.oracle import Oracle
    return Oracle(cloud, **oracle_params(params)).setup()


def test_c1_full_length_fp64():
    """9,600 points, first order, all 1000 iterations: every prim within 1e-10
    relative of the oracle's, the residue history within 1e-8 (a log10) and Cd within
    1e-10."""
    from mfcfd_amd import workloads as W
    cloud, p = W.cloud_of(W.C1), W.C1["params"]
    n_it = W.C1["iterations"]
    print(f"[c1] oracle {n_it} iterations on {cloud.n} points", flush=True)
    o = _oracle(cloud, p)
    res_o, _ = o.run(n_it)
    o.compute_cl_cd_cm()
    with _gpu(cloud, p) as g:
        res_g, _, cd_g = g.q_lskum(n_it)
        ep = rel_err(g.prim, o.field("prim"))
        print(f"[c1] prim rel err after {n_it} iterations {ep:.3e}, final residue "
              f"{res_g[-1]:.4f}", flush=True)
        assert len(res_g) == n_it == len(res_o)
        assert ep < 1e-10
        assert np.abs(res_g - res_o).max() < 1e-8
        assert abs(cd_g[-1] - o.field("Cd")[0]) <= 1e-10 * abs(o.field("Cd")[0])


@pytest.mark.timeout(600)
def test_c2_fp64_against_oracle():
    """160,000 points, 50 FP64 iterations under the autotuned shapes: every prim
    within 1e-10 relative of the oracle's, residues within 1e-8, Cl and Cd within
    1e-10."""
    from mfcfd_amd import workloads as W
    cloud, p = W.cloud_of(W.C2), W.C2["params"]
    n_it = 50
    print(f"[c2] oracle {n_it} iterations on {cloud.n} points", flush=True)
    o = _oracle(cloud, p)
    res_o, _ = o.run(n_it)
    o.compute_cl_cd_cm()
    with _gpu(cloud, p) as g:
        g.autotune()
        res_g, _, _ = g.q_lskum(n_it)
        cl, cd, _, _ = g.compute_cl_cd_cm()
        ep = rel_err(g.prim, o.field("prim"))
        print(f"[c2] prim rel err after {n_it} iterations {ep:.3e}", flush=True)
        assert ep < 1e-10
        assert np.abs(res_g - res_o).max() < 1e-8
        assert abs(cl[0] - o.field("Cl")[0]) <= 1e-10 * abs(o.field("Cl")[0])
        assert abs(cd[0] - o.field("Cd")[0]) <= 1e-10 * abs(o.field("Cd")[0])


def test_c2_mixed_full_length():
    """160,000 points, all 5000 iterations: the mixed path under the autotuned shapes
    within |dCl|, |dCd| <= 1e-4 of the FP64 path, and the two converge alike (final
    residues within 0.5 of a decade)."""
    from mfcfd_amd import workloads as W
    cloud, p = W.cloud_of(W.C2), W.C2["params"]
    n_it = W.C2["iterations"]
    with _gpu(cloud, p) as a, _gpu(cloud, p, precision="mixed") as b:
        a.autotune()
        b.autotune()
        ra, _, _ = a.q_lskum(n_it)
        rb, _, _ = b.q_lskum(n_it)
        cla, cda, _, _ = a.compute_cl_cd_cm()
        clb, cdb, _, _ = b.compute_cl_cd_cm()
        dp = rel_err(b.prim, a.prim)
        print(f"[c2] {n_it} iterations: Cl {cla[0]:.8f} Cd {cda[0]:.8f}; mixed dCl "
              f"{abs(cla[0] - clb[0]):.2e} dCd {abs(cda[0] - cdb[0]):.2e}; prim rel "
              f"{dp:.2e}; residue fp64 {ra[-1]:.3f} mixed {rb[-1]:.3f}", flush=True)
        assert len(ra) == len(rb) == n_it
        assert abs(cla[0] - clb[0]) <= 1e-4 and abs(cda[0] - cdb[0]) <= 1e-4
        assert abs(ra[-1] - rb[-1]) < 0.5


def test_c2_staged_bits():
    """Staged walks give the global walks' bits at 160,000 points (mixed path,
    5 iterations, 4 lanes per point)."""
    from mfcfd_amd import workloads as W
    cloud, p = W.cloud_of(W.C2), W.C2["params"]
    with _gpu(cloud, p, precision="mixed") as s, _gpu(cloud, p, precision="mixed") as t:
        s.tuning = {"lpp_lsq": 16 | 4, "lpp_inner": 16 | 4, "lpp_flux": 16 | 4}
        t.tuning = {"lpp_lsq": 4, "lpp_inner": 4, "lpp_flux": 4}
        s.q_lskum(5)
        t.q_lskum(5)
        assert np.array_equal(s.prim, t.prim)


@pytest.mark.timeout(900)
def test_c4_fp64_against_oracle_and_mixed_600():
    """2,560,000 points: 2 FP64 iterations against the oracle (every prim within
    1e-10 relative); then the FP64 and mixed paths over the workload's 600 iterations
    (100 + 500), |dCl|, |dCd| <= 1e-4 at the end."""
    from mfcfd_amd import workloads as W
    cloud, p = W.cloud_of(W.C4), W.C4["params"]
    print(f"[c4] oracle 2 iterations on {cloud.n} points", flush=True)
    o = _oracle(cloud, p)
    res_o, _ = o.run(2)
    o.compute_cl_cd_cm()
    ref = o.field("prim").copy()
    cd_o = float(o.field("Cd")[0])
    del o
    n_it = W.C4["warmup"] + W.C4["steps"]
    with _gpu(cloud, p) as a:
        a.autotune()
        res_g, _, cd_g = a.q_lskum(2)
        ep = rel_err(a.prim, ref)
        print(f"[c4] FP64 prim rel err after 2 iterations {ep:.3e}", flush=True)
        assert ep < 1e-10
        assert abs(cd_g[-1] - cd_o) <= 1e-10 * abs(cd_o)
        a.q_lskum(n_it - 2, first_it=3)
        cla, cda, _, _ = a.compute_cl_cd_cm()
    with _gpu(cloud, p, precision="mixed") as b:
        b.autotune()
        rb, _, _ = b.q_lskum(n_it)
        clb, cdb, _, _ = b.compute_cl_cd_cm()
    print(f"[c4] {n_it} iterations: Cl {cla[0]:.3e} Cd {cda[0]:.8f}; mixed dCl "
          f"{abs(cla[0] - clb[0]):.2e} dCd {abs(cda[0] - cdb[0]):.2e}", flush=True)
    assert len(rb) == n_it and np.isfinite(rb).all()
    assert abs(cla[0] - clb[0]) <= 1e-4 and abs(cda[0] - cdb[0]) <= 1e-4


@pytest.fixture(scope="module")
def c5():
    from mfcfd_amd import workloads as W
    return W.c5_cases()[0], W.C5["params"]


@pytest.mark.timeout(900)
def test_c5_batch_cases_match_single_contexts(c5):
    """All 64 batched cases, 50 FP64 iterations: each case's prim, residue history and
    Cd equal its own single-case context's (same launch shapes, so to rounding)."""
    from mfcfd_amd.solver import LskumBatch, LskumSolver
    cases, p = c5
    iters = 50
    with LskumBatch(cases, **p) as bt:
        res, _, cd = bt.run(iters)
        prim = bt.prim
        print(f"[c5] batch of {bt.ncases} cases, {bt.n} points", flush=True)
        for k, (cloud, mach, aoa) in enumerate(cases):
            with LskumSolver(cloud, mach=mach, aoa=aoa, **p) as s:
                r, _, c = s.q_lskum(iters)
                assert rel_err(prim[bt.case_slice(k)], s.prim) <= 1e-14, k
                assert np.abs(res[:, k] - r).max() <= 1e-12, k
                assert np.abs(cd[:, k, 0] - c).max() <= 1e-13 * np.abs(c).max(), k


@pytest.mark.timeout(600)
def test_c5_extreme_clouds_against_oracle(c5):
    """The batch's cases on the smallest and the largest cloud against the oracle
    after 2 FP64 iterations (every prim within 1e-10 relative)."""
    from mfcfd_amd.solver import LskumBatch
    cases, p = c5
    sizes = [c.n for c, _, _ in cases]
    picks = [int(np.argmin(sizes)), int(np.argmax(sizes))]
    with LskumBatch(cases, **p) as bt:
        bt.run(2)
        prim = bt.prim
        for k in picks:
            cloud, mach, aoa = cases[k]
            q = dict(p, mach=mach, aoa=aoa)
            print(f"[c5] case {k}: oracle 2 iterations on {cloud.n} points", flush=True)
            o = _oracle(cloud, q)
            o.run(2)
            ep = rel_err(prim[bt.case_slice(k)], o.field("prim"))
            print(f"[c5] case {k} prim rel err {ep:.3e}", flush=True)
            assert ep < 1e-10, k


@pytest.mark.timeout(900)
def test_c5_mixed_batch_full_length(c5):
    """All 64 cases, 1000 iterations under the autotuned shapes: per case
    |dCl|, |dCd| <= 1e-4 between the mixed and the FP64 batch."""
    from mfcfd_amd import workloads as W
    from mfcfd_amd.solver import LskumBatch
    cases, p = c5
    iters = W.C5["iterations"]
    with LskumBatch(cases, **p) as a, LskumBatch(cases, precision="mixed", **p) as b:
        a.autotune()
        b.autotune()
        ra, _, _ = a.run(iters)
        rb, _, _ = b.run(iters)
        cla, cda, _, _ = a.compute_cl_cd_cm()
        clb, cdb, _, _ = b.compute_cl_cd_cm()
    dcl = np.abs(cla[:, 0] - clb[:, 0])
    dcd = np.abs(cda[:, 0] - cdb[:, 0])
    print(f"[c5] {iters} iterations, 64 cases: max dCl {dcl.max():.2e} max dCd {dcd.max():.2e} "
          f"(median {np.median(dcd):.2e})", flush=True)
    assert np.isfinite(ra).all() and np.isfinite(rb).all()
    assert dcl.max() <= 1e-4 and dcd.max() <= 1e-4
