"""GPU parity: the MI355X path against the FP64 oracle on the same clouds and states.

Every test goes through the C ABI (libmfcfd_hip.so).  The oracle (oracle/) is the
checker: it restates src_mpi_primal in C++ and is itself pinned against the
reference's own compiled routines and closed forms (tests/test_oracle.py).

Tolerances (relative to the field's max magnitude):
  * single routines in FP64: 1e-12 (q, dq, qm, delta), 1e-10 (flux_res: an LSQ
    derivative of differences of nearly equal fluxes, where libm-vs-device
    transcendentals at the ulp level are amplified by cancellation);
  * FP64 trajectories: per-point prim within 1e-10 after the iterations run, residue
    history within 1e-8 absolute (it is a log10), Cd within 1e-10 relative.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

PARAMS = dict(cfl=0.2, mach=0.63, aoa=2.0, vl_const=5.0, inner_iterations=2)


def _oracle(cloud, **kw):
    from oracle.oracle import Oracle
    p = dict(PARAMS)
    p.update(kw)
    p["aoa_deg"] = p.pop("aoa")
    return Oracle(cloud, **p).setup()


def _gpu(cloud, **kw):
    from mfcfd_amd.solver import LskumSolver
    p = dict(PARAMS)
    p.update(kw)
    return LskumSolver(cloud, **p)


@pytest.fixture(scope="module")
def developed(small_cloud):
    """Oracle state after 40 FP64 iterations from freestream (limiter active)."""
    o = _oracle(small_cloud)
    print(f"[developed] oracle 40 iterations on {small_cloud.n} points", flush=True)
    o.run(40)
    return o


def test_normals_match(small_cloud):
    o = _oracle(small_cloud)
    with _gpu(small_cloud) as g:
        assert rel_err(g.field("nx"), o.field("nx")) == 0.0
        assert rel_err(g.field("ny"), o.field("ny")) == 0.0


def test_q_variables(small_cloud, developed):
    o = developed
    with _gpu(small_cloud) as g:
        g.prim = o.field("prim")
        g.eval_q_variables()
        o.eval_q_variables()
        assert rel_err(g.field("q"), o.field("q")) < 1e-14


def test_q_derivatives_and_inner_loop(small_cloud, developed):
    o = developed
    with _gpu(small_cloud) as g:
        g.prim = o.field("prim")
        o.eval_q_variables()
        g.eval_q_derivatives()
        o.eval_q_derivatives()
        assert rel_err(g.field("qm"), o.field("qm")) < 1e-15  # q from device log: ulp level
        assert rel_err(g.field("dq"), o.field("dq")) < 1e-12
        for _ in range(3):
            g.eval_q_inner_loop()
            o.eval_q_inner_loop()
            o.eval_update_innerloop()
            assert rel_err(g.field("dq"), o.field("dq")) < 1e-12


def test_func_delta(small_cloud, developed):
    o = developed
    with _gpu(small_cloud) as g:
        g.prim = o.field("prim")
        g.func_delta()
        o.func_delta()
        assert rel_err(g.field("delta"), o.field("delta")) < 1e-14


def test_flux_residual_same_inputs(small_cloud, developed):
    """cal_flux_residual on identical q, dq, qm, delta (oracle's, copied in)."""
    o = developed
    o.eval_q_variables()
    o.eval_q_derivatives()
    o.eval_q_inner_loop()
    o.eval_update_innerloop()
    o.func_delta()
    o.cal_flux_residual()
    with _gpu(small_cloud) as g:
        g.prim = o.field("prim")
        g.set_field("q", o.field("q"))
        g.set_field("dq", o.field("dq"))
        g.set_field("qm", o.field("qm"))
        g.set_field("delta", o.field("delta"))
        g.cal_flux_residual()
        e = rel_err(g.field("flux_res"), o.field("flux_res"))
        print(f"flux_res rel err {e:.3e}")
        assert e < 1e-10


@pytest.mark.parametrize("rk", [1, 2, 3, 4])
def test_state_update(small_cloud, developed, rk):
    """One full stage through the reference routines on both sides."""
    o = developed
    prim0 = o.field("prim").copy()
    with _gpu(small_cloud) as g:
        g.prim = prim0
        g.set_field("prim_old", prim0 * (1.0 + 1e-3))
        o.field("prim_old")[:] = prim0 * (1.0 + 1e-3)
        for s in (g, o):
            s.func_delta()
            s.eval_q_variables()
            s.eval_q_derivatives()
        g.cal_flux_residual()
        o.cal_flux_residual()
        sg = g.state_update(rk)
        o.state_update(rk)
        so = o.scalar("sum_res_sqr")
        assert rel_err(g.prim, o.field("prim")) < 1e-12
        assert abs(sg - so) <= 1e-10 * abs(so)
        o.field("prim")[:] = prim0


@pytest.mark.parametrize("inner,tscheme", [(0, "ssprk43"), (2, "ssprk43"), (1, "first")])
def test_iterations_fp64(small_cloud, inner, tscheme):
    """q_lskum loop: 30 fused, graph-replayed iterations from freestream."""
    o = _oracle(small_cloud, inner_iterations=inner,
                rks=4 if tscheme == "ssprk43" else 1, euler=1.0 if tscheme == "ssprk43" else 2.0)
    res_o, rn_o = o.run(30)
    o.compute_cl_cd_cm()
    with _gpu(small_cloud, inner_iterations=inner, tscheme=tscheme) as g:
        res_g, rn_g, cd_g = g.q_lskum(30)
        assert len(res_g) == 30
        ep = rel_err(g.prim, o.field("prim"))
        print(f"prim rel err after 30 its: {ep:.3e}")
        assert ep < 1e-10
        assert np.abs(res_g - res_o).max() < 1e-8
        assert rel_err(rn_g, rn_o) < 1e-10
        cl, cd, cm, cp = g.compute_cl_cd_cm()
        assert abs(cd[0] - o.field("Cd")[0]) <= 1e-10 * max(abs(o.field("Cd")[0]), 1e-12)
        assert abs(cl[0] - o.field("Cl")[0]) <= 1e-10 * max(abs(o.field("Cl")[0]), 1e-12)
        assert rel_err(cp, o.field("cp")) < 1e-10
        assert abs(cd_g[-1] - cd[0]) <= 1e-14 * max(abs(cd[0]), 1e-300)


@pytest.mark.parametrize("tscheme", ["ssprk43", "first"])
def test_graphs_built_at_tuning(small_cloud, tscheme):
    """mf_set_tuning builds the replayed graphs (16 iterations and 1, from both buffer
    phases when a stage count is odd) before the next run: 17 and then 3 iterations
    after a tuning give the bits of a context whose graphs are built inside its runs."""
    with _gpu(small_cloud, tscheme=tscheme) as a, _gpu(small_cloud, tscheme=tscheme) as b:
        a.tuning = b.tuning  # the same shapes, set on a only
        ra1, _, _ = a.q_lskum(17, first_it=1)
        ra2, _, _ = a.q_lskum(3, first_it=18)
        rb, _, _ = b.q_lskum(20, first_it=1)
        assert np.array_equal(a.prim, b.prim)
        assert np.array_equal(np.concatenate([ra1, ra2]), rb)


def test_split_runs_continue(small_cloud):
    """Two mf_run calls (10 + 10 iterations) equal one of 20, residue included."""
    with _gpu(small_cloud) as a, _gpu(small_cloud) as b:
        r1, _, _ = a.q_lskum(10, first_it=1)
        r2, _, _ = a.q_lskum(10, first_it=11)
        rb, _, _ = b.q_lskum(20, first_it=1)
        assert np.array_equal(np.concatenate([r1, r2]), rb)
        assert np.array_equal(a.prim, b.prim)


def test_reorder_invariance(small_cloud):
    """Hilbert renumbering changes only the memory order, not the arithmetic."""
    with _gpu(small_cloud, reorder=1) as a, _gpu(small_cloud, reorder=0) as b:
        a.q_lskum(15)
        b.q_lskum(15)
        assert rel_err(a.prim, b.prim) < 1e-13


@pytest.mark.parametrize("lpp", [2, 4, 8])
def test_lanes_per_point_shapes(small_cloud, lpp):
    """Splitting each point's stencil over 2/4/8 lanes only reorders sums: the
    trajectory stays within 1e-12 of the one-lane run and of the oracle's tolerance."""
    o = _oracle(small_cloud)
    o.run(20)
    with _gpu(small_cloud) as a, _gpu(small_cloud) as b:
        b.tuning = {"lpp_lsq": lpp, "lpp_inner": lpp, "lpp_flux": lpp}
        assert {k: b.tuning[k] for k in ("lpp_lsq", "lpp_inner", "lpp_flux")} == \
            {"lpp_lsq": lpp, "lpp_inner": lpp, "lpp_flux": lpp}
        ra, _, cda = a.q_lskum(20)
        rb, _, cdb = b.q_lskum(20)
        assert rel_err(b.prim, a.prim) < 1e-12
        assert np.abs(ra - rb).max() < 1e-9
        assert abs(cda[-1] - cdb[-1]) <= 1e-12 * abs(cda[-1])
        assert rel_err(b.prim, o.field("prim")) < 1e-10


def test_autotune_and_invalid_shape(small_cloud):
    """mf_autotune picks a valid shape per kernel and leaves the state untouched;
    mf_set_tuning rejects shapes other than 1, 2, 4, 8."""
    from mfcfd_amd._lib import MfError
    with _gpu(small_cloud) as a, _gpu(small_cloud) as b:
        a.q_lskum(5)
        before = a.prim.copy()
        t = a.autotune(reps=3)
        staged = tuple(16 | v for v in (1, 2, 4))
        resident = tuple(32 | v for v in (1, 2, 4))
        assert t["lpp_lsq"] in (1, 2, 4, 8) + staged
        assert t["lpp_inner"] in (1, 2, 4, 8) + staged + resident
        assert t["lpp_flux"] in (1, 2, 4, 8, 16 | 2, 16 | 4, 16 | 8)
        assert all(t[k] > 0 for k in ("us_lsq", "us_inner", "us_flux"))
        assert np.array_equal(a.prim, before)
        # the tuned context continues the same trajectory as an untuned one
        b.q_lskum(5)
        ra, _, _ = a.q_lskum(5, first_it=6)
        rb, _, _ = b.q_lskum(5, first_it=6)
        assert rel_err(a.prim, b.prim) < 1e-12
        with pytest.raises(MfError):
            a.tuning = {"lpp_flux": 3}
        with pytest.raises(MfError):
            a.tuning = {"lpp_flux": 16 | 1}  # staged flux: 2, 4 or 8 lanes
        with pytest.raises(MfError):
            a.tuning = {"lpp_inner": 16 | 8}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lpp", [1, 2, 4])
def test_staged_walk_bit_identical(c1_cloud, precision, lpp):
    """The staged walks (neighbour records copied once per block into LDS) keep each
    lane's slots and summation order: LSQ, inner and flux+update kernels give the same
    bits as the global walks with the same lanes per point."""
    kw = dict(precision=precision, inner_iterations=2)
    lf = 2 * lpp  # flux kernel: 2, 4, 8 lanes (32-point blocks when staged)
    with _gpu(c1_cloud, **kw) as a, _gpu(c1_cloud, **kw) as b:
        a.tuning = {"lpp_lsq": lpp, "lpp_inner": lpp, "lpp_flux": lf}
        b.tuning = {"lpp_lsq": 16 | lpp, "lpp_inner": 16 | lpp, "lpp_flux": 16 | lf}
        assert b.tuning == {**b.tuning, "lpp_lsq": 16 | lpp, "lpp_inner": 16 | lpp, "lpp_flux": 16 | lf}
        ra, _, cda = a.q_lskum(25)
        rb, _, cdb = b.q_lskum(25)
        assert np.array_equal(a.prim, b.prim)
        assert np.array_equal(cda, cdb)
        # the residual norm's partial sums are grouped per block, and the staged flux
        # kernel's blocks hold 32 points: same terms, another summation grouping
        assert np.abs(ra - rb).max() <= 1e-12


def test_nan_residue_stops_loop(small_cloud):
    """q_lskum.F90:71 exits after the first NaN residue; so does mf_run."""
    o = _oracle(small_cloud)
    prim = o.field("prim")
    prim[100, 3] = -0.5  # negative pressure: log of a negative number
    res_o, _ = o.run(10)
    with _gpu(small_cloud) as g:
        p = g.prim
        p[100, 3] = -0.5
        g.prim = p
        res_g, _, _ = g.q_lskum(10)
    assert len(res_g) == len(res_o)
    assert np.isnan(res_g[-1]) and np.isnan(res_o[-1])


def test_second_order_vl150_fp64(c1_cloud):
    """Same cloud, second order with the default limiter constant (VL_CONST 150):
    100 iterations agree."""
    o = _oracle(c1_cloud, inner_iterations=0, vl_const=150.0)
    res_o, _ = o.run(100)
    with _gpu(c1_cloud, inner_iterations=0, vl_const=150.0) as g:
        res_g, _, _ = g.q_lskum(100)
        assert rel_err(g.prim, o.field("prim")) < 1e-10
        assert np.abs(res_g - res_o).max() < 1e-8


def test_mixed_precision_close(small_cloud):
    """Mixed path (FP32 state/flux/limiter/update, FP64 LSQ and sums) stays close to
    FP64 after 60 iterations: |dCl|, |dCd| <= 1e-4 and prim within 1e-4."""
    with _gpu(small_cloud) as a, _gpu(small_cloud, precision="mixed") as b:
        a.q_lskum(60)
        b.q_lskum(60)
        cla, cda, _, _ = a.compute_cl_cd_cm()
        clb, cdb, _, _ = b.compute_cl_cd_cm()
        print(f"mixed: dCl {abs(cla[0]-clb[0]):.2e} dCd {abs(cda[0]-cdb[0]):.2e} "
              f"prim {rel_err(b.prim, a.prim):.2e}")
        assert abs(cla[0] - clb[0]) <= 1e-4
        assert abs(cda[0] - cdb[0]) <= 1e-4
        assert rel_err(b.prim, a.prim) < 1e-4


def test_profile_kernels_runs(small_cloud):
    with _gpu(small_cloud) as g:
        t = g.profile_kernels(3)
        assert t["flux_update"][0] > 0 and t["flux_update"][1] == 4
        assert t["lsq"][1] == 4 and t["inner"][1] == 8


@pytest.mark.parametrize("power", [-2.0, 1.0])
def test_distance_weights_power(small_cloud, power):
    """-power p weights each edge of the least-squares sums and of the flux derivatives by
    dist**p (q_variables.F90:91,169; interior_fluxes.F90:71,170,270,368): a 20-iteration
    trajectory against the oracle, and the staged walks against the global ones."""
    o = _oracle(small_cloud, power=power)
    o.run(20)
    with _gpu(small_cloud, power=power) as a, _gpu(small_cloud, power=power) as b:
        b.tuning = {"lpp_lsq": 16 | 2, "lpp_inner": 16 | 2, "lpp_flux": 16 | 4}
        a.tuning = {"lpp_lsq": 2, "lpp_inner": 2, "lpp_flux": 4}
        ra, _, cda = a.q_lskum(20)
        b.q_lskum(20)
        assert rel_err(a.prim, o.field("prim")) < 1e-10
        assert np.array_equal(a.prim, b.prim)
    with _gpu(small_cloud, power=power) as c:
        c.q_lskum(20)
        assert rel_err(c.prim, o.field("prim")) < 1e-10


@pytest.mark.parametrize("power", [-2.0, 1.0])
def test_mixed_distance_weights(small_cloud, power):
    """The mixed path's distance weights (FP32 edge geometry, weight from the FP64
    distance): after 60 iterations |dCl|, |dCd| <= 1e-4 of the FP64 path at the same
    -power, and its staged walks give the global walks' bits."""
    kw = dict(power=power)
    with _gpu(small_cloud, **kw) as a, _gpu(small_cloud, precision="mixed", **kw) as b, \
            _gpu(small_cloud, precision="mixed", **kw) as c:
        b.tuning = {"lpp_lsq": 2, "lpp_inner": 2, "lpp_flux": 4}
        c.tuning = {"lpp_lsq": 16 | 2, "lpp_inner": 16 | 2, "lpp_flux": 16 | 4}
        a.q_lskum(60)
        b.q_lskum(60)
        c.q_lskum(60)
        cla, cda, _, _ = a.compute_cl_cd_cm()
        clb, cdb, _, _ = b.compute_cl_cd_cm()
        assert abs(cla[0] - clb[0]) <= 1e-4 and abs(cda[0] - cdb[0]) <= 1e-4
        assert rel_err(b.prim, a.prim) < 1e-4
        assert np.array_equal(b.prim, c.prim)


@pytest.mark.parametrize("kw", [dict(first_order=1, inner_iterations=0), dict(tscheme="first", inner_iterations=1)],
                         ids=["first_order", "tscheme_first"])
def test_mixed_other_schemes(small_cloud, kw):
    """The mixed path with q-tilde = q (first order) and with the one-stage time scheme
    (rks 1, euler 2): after 60 iterations |dCl|, |dCd| <= 1e-4 of the FP64 path."""
    with _gpu(small_cloud, **kw) as a, _gpu(small_cloud, precision="mixed", **kw) as b:
        a.q_lskum(60)
        b.q_lskum(60)
        cla, cda, _, _ = a.compute_cl_cd_cm()
        clb, cdb, _, _ = b.compute_cl_cd_cm()
        assert abs(cla[0] - clb[0]) <= 1e-4 and abs(cda[0] - cdb[0]) <= 1e-4
        assert rel_err(b.prim, a.prim) < 1e-4


def _relabelled(cloud, **fields):
    import dataclasses
    return dataclasses.replace(cloud, **{k: v.copy() for k, v in fields.items()})


@pytest.mark.parametrize("case", ["normal_flag_1", "format_2"])
def test_interior_normal_options(small_cloud, case):
    """point_normals.F90:90-103: interior normals (1, 0) with
    -interior_points_normal_flag 1, and the file's normals with -format_file quadtree
    (format 2; here seeded random unit vectors), which put every interior point's
    split fluxes in a rotated frame.  20 iterations against the oracle."""
    c = small_cloud
    kw = {}
    if case == "normal_flag_1":
        kw["interior_points_normal_flag"] = 1
    else:
        rng = np.random.default_rng(7)
        ang = rng.uniform(0.0, 2.0 * np.pi, c.n)
        c = _relabelled(c, nx=np.cos(ang), ny=np.sin(ang))
        kw["format"] = 2
    o = _oracle(c, **kw)
    o.run(20)
    o.compute_cl_cd_cm()
    with _gpu(c, **kw) as g:
        # A 90-degree turn of the interior frame only reassigns the split stencils, so
        # normal flag 1 leaves the trajectory unchanged up to rounding: check the
        # normals themselves as well.
        assert np.array_equal(g.field("nx"), o.field("nx"))
        assert np.array_equal(g.field("ny"), o.field("ny"))
        interior = c.flag_1 == 1
        if case == "normal_flag_1":
            assert np.all(g.field("nx")[interior] == 1.0) and np.all(g.field("ny")[interior] == 0.0)
        else:
            assert np.array_equal(g.field("nx")[interior], c.nx[interior])
        g.q_lskum(20)
        assert rel_err(g.prim, o.field("prim")) < 1e-10
        _, cd, _, _ = g.compute_cl_cd_cm()
        assert abs(cd[0] - o.field("Cd")[0]) <= 1e-10 * abs(o.field("Cd")[0])


@pytest.mark.parametrize("case", ["normal_flag_1", "format_2"])
def test_mixed_rotated_interior_frame(small_cloud, case):
    """The mixed path's interior points in a rotated frame (normals (1, 0), or the file's
    normals): the general form of the flux kernel, not the (x, y) frame it takes when
    every interior normal is (0, 1).  Against the FP64 path after 60 iterations:
    |dCl|, |dCd| <= 1e-4 and every prim within 1e-4 relative."""
    c = small_cloud
    kw = {}
    if case == "normal_flag_1":
        kw["interior_points_normal_flag"] = 1
    else:
        rng = np.random.default_rng(7)
        ang = rng.uniform(0.0, 2.0 * np.pi, c.n)
        c = _relabelled(c, nx=np.cos(ang), ny=np.sin(ang))
        kw["format"] = 2
    with _gpu(c, **kw) as a, _gpu(c, precision="mixed", **kw) as b:
        a.q_lskum(60)
        b.q_lskum(60)
        cla, cda, _, _ = a.compute_cl_cd_cm()
        clb, cdb, _, _ = b.compute_cl_cd_cm()
        assert abs(cla[0] - clb[0]) <= 1e-4 and abs(cda[0] - cdb[0]) <= 1e-4
        assert rel_err(b.prim, a.prim) < 1e-4


def test_two_shapes(small_cloud):
    """-shapes 2 with the lower surface's wall points relabelled flag_2 = 2: per-shape
    Cl, Cd, Cm (compute_force_coeffs.F90:11-102) against the oracle, and the two shapes'
    sums equal the one-shape forces (the integral is split, not changed)."""
    c = small_cloud
    f2 = c.flag_2.copy()
    lower = (c.flag_1 == 0) & (c.y < 0.0)
    f2[lower] = 2
    c2 = _relabelled(c, flag_2=f2)
    o = _oracle(c2, shapes=2)
    o.run(15)
    o.compute_cl_cd_cm()
    with _gpu(c2, shapes=2) as g, _gpu(c) as h:
        g.q_lskum(15)
        h.q_lskum(15)
        assert np.array_equal(g.prim, h.prim)  # the shapes only label force sums
        cl, cd, cm, _ = g.compute_cl_cd_cm()
        cl1, cd1, cm1, _ = h.compute_cl_cd_cm()
        assert cl.shape == (2,) and cl1.shape == (1,)
        for got, ref in ((cl, o.field("Cl")), (cd, o.field("Cd")), (cm, o.field("Cm"))):
            assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
        for two, one in ((cl, cl1), (cd, cd1), (cm, cm1)):
            assert abs(two.sum() - one[0]) <= 1e-12 * abs(one[0]) + 1e-15


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_time_kernels_keeps_state(small_cloud, precision):
    """mf_time_kernels launches the LSQ, inner and flux+update kernels at the current
    shapes on the current state and writes only what an iteration recomputes: the
    iterations after it give the same bits as without it."""
    kw = dict(precision=precision)
    with _gpu(small_cloud, **kw) as a, _gpu(small_cloud, **kw) as b:
        a.q_lskum(5)
        b.q_lskum(5)
        us = b.time_kernels(5)
        assert all(v > 0.0 for v in us.values())
        ra, _, cda = a.q_lskum(10, first_it=6)
        rb, _, cdb = b.q_lskum(10, first_it=6)
        assert np.array_equal(a.prim, b.prim)
        assert np.array_equal(ra, rb) and np.array_equal(cda, cdb)
