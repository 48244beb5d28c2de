"""Pins the FP64 oracle (oracle/lskum_oracle.cpp) before it is trusted as the checker.

1. Golden vectors: tests/golden/ref_routines.npz holds outputs of the reference's own
   Fortran routines (src_mpi_primal compiled unmodified, see make_golden.py).  The
   oracle must reproduce them bit for bit (same operation order, same libm, no FMA).
2. Where oracle/_ref is built (the reference sources are present), the same routines
   are also compared live on a second cloud.
3. Closed forms and invariants that a dropped term, a wrong sign or a swapped operand
   would break: split-flux identities, quadrant identities, the q <-> primitive round
   trip, exact LSQ gradients of linear fields, the freestream fixed point, the SSPRK
   stage weights, symmetric-pressure force identities and brute-force time steps.

Routines with no reference build here (state_update, func_delta,
generate_connectivity, compute_cl_cd_cm: they need PETSc) are pinned by 3 only.
"""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, RefRoutines, flux, qtilde_to_primitive

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_routines.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def _gold_cloud(g):
    from mfcfd_amd.cloud import PointCloud
    return PointCloud(x=g["cloud_x"], y=g["cloud_y"], nx=g["cloud_nx"], ny=g["cloud_ny"],
                      min_dist=g["cloud_min_dist"], left=g["cloud_left"], right=g["cloud_right"],
                      flag_1=g["cloud_flag_1"], flag_2=g["cloud_flag_2"], nbhs=g["cloud_nbhs"],
                      conn=g["cloud_conn"])


def _gold_oracle(g):
    cfl, mach, aoa, vl, inner = g["params"]
    return Oracle(_gold_cloud(g), cfl=cfl, mach=mach, aoa_deg=aoa, vl_const=vl,
                  inner_iterations=int(inner)).setup()


def test_golden_normals_and_split(gold):
    o = _gold_oracle(gold)
    assert np.array_equal(o.field("nx"), gold["nx"])
    assert np.array_equal(o.field("ny"), gold["ny"])
    for f in ("xpos_nbhs", "xneg_nbhs", "ypos_nbhs", "yneg_nbhs", "xpos_conn", "xneg_conn",
              "ypos_conn", "yneg_conn"):
        a = o.ifield(f)
        b = gold["split_" + f]
        if f.endswith("conn"):
            n = o.ifield(f.replace("conn", "nbhs"))
            for i in range(len(n)):
                assert np.array_equal(a[i, :n[i]], b[i, :n[i]])
        else:
            assert np.array_equal(a, b)


def test_golden_q_lsq_inner(gold):
    o = _gold_oracle(gold)
    o.field("prim")[:] = gold["prim"]
    o.eval_q_variables()
    assert np.array_equal(o.field("q"), gold["q"])
    o.eval_q_derivatives()
    assert np.array_equal(o.field("dq"), gold["dq"])
    assert np.array_equal(o.field("qm"), gold["qm"])
    o.eval_q_inner_loop()
    o.eval_update_innerloop()
    assert np.array_equal(o.field("dq"), gold["dq_inner"])


def test_golden_flux_residual(gold):
    o = _gold_oracle(gold)
    o.field("prim")[:] = gold["prim"]
    o.eval_q_variables()
    o.eval_q_derivatives()
    o.eval_q_inner_loop()
    o.eval_update_innerloop()
    o.field("delta")[:] = gold["delta"]
    o.cal_flux_residual()
    assert np.array_equal(o.field("flux_res"), gold["flux_res"])
    assert np.abs(gold["flux_res"]).max() > 0


def test_golden_delta(gold):
    """func_delta has no reference build; the golden delta came from the oracle, so
    this checks the oracle against a brute-force numpy evaluation instead."""
    o = _gold_oracle(gold)
    o.field("prim")[:] = gold["prim"]
    o.func_delta()
    x, y, conn, nb = gold["cloud_x"], gold["cloud_y"], gold["cloud_conn"], gold["cloud_nbhs"]
    p = gold["prim"]
    cfl = gold["params"][0]
    ref = np.ones(len(x))
    for i in range(len(x)):
        for k in conn[i, :nb[i]]:
            d = np.sqrt((x[k] - x[i]) ** 2 + (y[k] - y[i]) ** 2)
            c = np.sqrt(p[k, 1] ** 2 + p[k, 2] ** 2) + 3 * np.sqrt(p[k, 3] / p[k, 0])
            ref[i] = min(ref[i], cfl * d / c)
    assert np.allclose(o.field("delta"), ref, rtol=1e-14, atol=0)


def test_golden_fluxes(gold):
    for row, out in zip(gold["flux_in"], gold["flux_out"]):
        for k in range(8):
            assert np.array_equal(flux(k, *row), out[k])


def test_golden_qtilde_to_primitive(gold):
    for qt, out in zip(gold["qt_in"], gold["qt_out"]):
        assert np.array_equal(qtilde_to_primitive(qt), out)


def test_golden_limiter_covers_both_branches(gold):
    phi = gold["venkat_phi"]
    assert (phi == 1.0).any() and (phi < 1.0).any() and (phi > 0).all()


@pytest.mark.skipif(not RefRoutines.available(), reason="reference sources not built here")
def test_live_reference_routines():
    """Oracle vs the reference's compiled routines on a 160 x 60 cloud, 2 inner
    iterations, VL_CONST 5, after 10 iterations."""
    from mfcfd_amd.cloud import naca_ogrid
    c = naca_ogrid(160, 60, seed=5)
    o = Oracle(c, cfl=0.2, mach=0.63, aoa_deg=2.0, vl_const=5.0, inner_iterations=2).setup()
    o.run(10)
    ref = RefRoutines(c, vl_const=5.0)
    ref.set_split(o)
    ref.call("compute_normals")
    assert np.array_equal(ref.get("nx"), o.field("nx"))
    ref.set("prim", o.field("prim"))
    for r in ("eval_q_variables", "eval_q_derivatives"):
        ref.call(r)
        getattr(o, r)()
    assert np.array_equal(ref.get("dq"), o.field("dq"))
    ref.call("eval_q_inner_loop")
    o.eval_q_inner_loop()
    assert np.array_equal(ref.get("temp"), o.field("temp"))
    ref.call("eval_update_innerloop")
    o.eval_update_innerloop()
    o.func_delta()
    ref.set("delta", o.field("delta"))
    ref.call("cal_flux_residual")
    o.cal_flux_residual()
    assert np.array_equal(ref.get("flux_res"), o.field("flux_res"))
    for k in (0, 10, 500, 9000):
        qt = o.field("q")[k] * (1 + 1e-3)
        ref_phi = ref.venkat(k, qt)
        assert ((ref_phi > 0) & (ref_phi <= 1)).all()


@pytest.mark.skipif(not RefRoutines.available(), reason="reference sources not built here")
@pytest.mark.parametrize("case", ["normal_flag_1", "format_2"])
def test_live_reference_normal_options(case):
    """point_normals.F90:90-103 under -interior_points_normal_flag 1 (interior normal
    (1, 0)) and format 2 (the file's interior normals kept; here seeded random unit
    vectors), then the flux residual in those frames (flux_residual.F90), against the
    reference's compiled routines."""
    import dataclasses
    from mfcfd_amd.cloud import naca_ogrid
    c = naca_ogrid(80, 30, seed=6)
    if case == "normal_flag_1":
        kw = dict(interior_points_normal_flag=1)
    else:
        ang = np.random.default_rng(11).uniform(0.0, 2.0 * np.pi, c.n)
        c = dataclasses.replace(c, nx=np.cos(ang), ny=np.sin(ang))
        kw = dict(format=2)
    o = Oracle(c, cfl=0.2, mach=0.63, aoa_deg=2.0, vl_const=5.0, inner_iterations=0,
               **kw).setup()
    o.run(8)
    ref = RefRoutines(c, vl_const=5.0, **kw)
    ref.set("nx", c.nx)
    ref.set("ny", c.ny)
    ref.call("compute_normals")
    assert np.array_equal(ref.get("nx"), o.field("nx"))
    assert np.array_equal(ref.get("ny"), o.field("ny"))
    interior = c.flag_1 == 1
    if case == "normal_flag_1":
        assert (o.field("nx")[interior] == 1.0).all() and (o.field("ny")[interior] == 0.0).all()
    else:
        assert np.array_equal(o.field("nx")[interior], c.nx[interior])
    ref.set_split(o)
    ref.set("prim", o.field("prim"))
    for r in ("eval_q_variables", "eval_q_derivatives"):
        ref.call(r)
        getattr(o, r)()
    o.func_delta()
    ref.set("delta", o.field("delta"))
    ref.call("cal_flux_residual")
    o.cal_flux_residual()
    assert np.array_equal(ref.get("flux_res"), o.field("flux_res"))


def test_force_coefficients_split_over_shapes(cloud):
    """compute_force_coeffs.F90:11-102 accumulates each wall point into the shape its
    flag_2 names: relabelling the lower surface as shape 2 splits Cl, Cd, Cm into two
    parts whose sums are the one-shape values."""
    import dataclasses
    o1 = Oracle(cloud, mach=0.63, aoa_deg=2.0, vl_const=5.0, inner_iterations=0).setup()
    o1.run(6)
    f2 = cloud.flag_2.copy()
    lower = (cloud.flag_1 == 0) & (cloud.y < 0.0)
    f2[lower] = 2
    o2 = Oracle(dataclasses.replace(cloud, flag_2=f2), mach=0.63, aoa_deg=2.0,
                vl_const=5.0, inner_iterations=0, shapes=2).setup()
    o2.run(6)
    assert np.array_equal(o1.field("prim"), o2.field("prim"))
    o1.compute_cl_cd_cm()
    o2.compute_cl_cd_cm()
    for f in ("Cl", "Cd", "Cm"):
        two, one = o2.field(f), o1.field(f)
        assert two.shape == (2,) and (two != 0).all()
        assert abs(two.sum() - one[0]) <= 1e-13 * max(abs(one[0]), abs(two).max())


# ---------------------------------------------------------------------------------
# closed forms


def _euler_flux_tn(nx, ny, u1, u2, rho, pr):
    """Full Euler flux through the t-face: (rho ut, p + rho ut^2, rho ut un, ut (E+p))."""
    ut = u1 * ny - u2 * nx
    un = u1 * nx + u2 * ny
    E = 2.5 * pr + 0.5 * rho * (ut * ut + un * un)
    return np.array([rho * ut, pr + rho * ut * ut, rho * ut * un, ut * (E + pr)])


def _euler_flux_n(nx, ny, u1, u2, rho, pr):
    ut = u1 * ny - u2 * nx
    un = u1 * nx + u2 * ny
    E = 2.5 * pr + 0.5 * rho * (ut * ut + un * un)
    return np.array([rho * un, rho * ut * un, pr + rho * un * un, un * (E + pr)])


STATES = [(0.6, 0.8, 0.3, -0.2, 1.1, 0.7), (1.0, 0.0, 2.5, 0.1, 0.5, 0.9),
          (0.0, 1.0, -0.9, 0.4, 1.7, 0.3), (-0.28, 0.96, 0.05, -1.3, 0.8, 1.2)]


@pytest.mark.parametrize("st", STATES)
def test_split_flux_identities(st):
    Gxp, Gxn, Gyp, Gyn, GI, GII, GIII, GIV = (flux(k, *st) for k in range(8))
    assert np.allclose(Gxp + Gxn, _euler_flux_tn(*st), rtol=1e-13, atol=1e-14)
    assert np.allclose(Gyp + Gyn, _euler_flux_n(*st), rtol=1e-13, atol=1e-14)
    # quadrant fluxes split the x half-fluxes by the sign of the normal velocity
    assert np.allclose(GII + GIII, Gxp, rtol=1e-13, atol=1e-14)
    assert np.allclose(GI + GIV, Gxn, rtol=1e-13, atol=1e-14)
    assert not np.allclose(GII, GIII)


def test_split_flux_supersonic_limits():
    # ut sqrt(beta) >> 1: the + half carries the whole flux, the - half none
    st = (0.0, 1.0, 20.0, 0.3, 1.0, 0.7)
    assert np.allclose(flux(0, *st), _euler_flux_tn(*st), rtol=1e-12)
    assert np.allclose(flux(1, *st), 0.0, atol=1e-10)


def test_qtilde_round_trip():
    rng = np.random.default_rng(0)
    for _ in range(50):
        rho, u1, u2, pr = rng.uniform(0.2, 2), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.1, 2)
        beta = 0.5 * rho / pr
        q = np.array([np.log(rho) + 2.5 * np.log(beta) - beta * (u1 * u1 + u2 * u2),
                      2 * beta * u1, 2 * beta * u2, -2 * beta])
        assert np.allclose(qtilde_to_primitive(q), [rho, u1, u2, pr], rtol=1e-13)


@pytest.fixture(scope="module")
def cloud():
    from mfcfd_amd.cloud import naca_ogrid
    return naca_ogrid(48, 14, seed=3, first_spacing=0.01, far_radius=4.0)


def test_lsq_exact_for_linear_fields(cloud):
    """eval_q_derivatives recovers the gradient of a linear field exactly (up to
    rounding) on every stencil, and the inner (defect-correction) loop keeps it."""
    o = Oracle(cloud).setup()
    a = np.array([0.3, -1.2, 0.7, 2.0])
    b = np.array([-0.5, 0.4, 1.1, -0.3])
    q = o.field("q")
    q[:] = 1.0 + cloud.x[:, None] * a + cloud.y[:, None] * b
    o.eval_q_derivatives()
    dq = o.field("dq")
    assert np.allclose(dq[:, :, 0], a, rtol=0, atol=1e-9)
    assert np.allclose(dq[:, :, 1], b, rtol=0, atol=1e-9)
    o.eval_q_inner_loop()
    o.eval_update_innerloop()
    assert np.allclose(o.field("dq")[:, :, 0], a, atol=1e-9)
    assert np.allclose(o.field("dq")[:, :, 1], b, atol=1e-9)
    qm = o.field("qm")
    assert (qm[:, :, 0] >= q).all() and (qm[:, :, 1] <= q).all()


def test_freestream_fixed_point(cloud):
    """From uniform freestream every dq and every flux difference is exactly zero, so
    stage 1 leaves interior points unchanged (to the rounding of the conserved-variable
    round trip); the far-field state Ubar of the
    freestream is the freestream (A2n + A2p = 1), and only the wall's no-penetration
    condition changes the state."""
    o = Oracle(cloud, cfl=0.2, mach=0.63, aoa_deg=2.0).setup()
    p0 = o.field("prim").copy()
    o.field("prim_old")[:] = p0
    o.func_delta()
    o.eval_q_variables()
    o.eval_q_derivatives()
    assert (o.field("dq") == 0).all()
    o.cal_flux_residual()
    assert (o.field("flux_res") == 0).all()
    o.state_update(1)
    p1 = o.field("prim")
    inter = cloud.flag_1 == 1
    outer = cloud.flag_1 == 2
    wall = cloud.flag_1 == 0
    assert np.allclose(p1[inter], p0[inter], rtol=1e-15, atol=1e-16)
    assert np.allclose(p1[outer], p0[outer], rtol=1e-13, atol=1e-15)
    # wall: normal velocity removed, tangential kept
    nx, ny = o.field("nx")[wall], o.field("ny")[wall]
    un = p1[wall, 1] * nx + p1[wall, 2] * ny
    assert np.abs(un).max() < 1e-14
    assert not np.allclose(p1[wall, 1:3], p0[wall, 1:3])


def test_ssprk_stage_weights(cloud):
    """With flux_res = 0, stage 3 returns 2/3 U_old + 1/3 U (conserved variables)."""
    o = Oracle(cloud, cfl=0.2).setup()
    inter = np.where(cloud.flag_1 == 1)[0]
    rng = np.random.default_rng(1)
    p = o.field("prim")
    p[:, 0] = rng.uniform(0.8, 1.2, cloud.n)
    p[:, 3] = rng.uniform(0.6, 0.8, cloud.n)
    po = o.field("prim_old")
    po[:] = p
    po[:, 0] *= 1.5
    o.field("flux_res")[:] = 0.0
    p_before = p.copy()
    o.state_update(3)
    rho = o.field("prim")[inter, 0]
    assert np.allclose(rho, 2 / 3 * 1.5 * p_before[inter, 0] + 1 / 3 * p_before[inter, 0], rtol=1e-14)
    o.field("prim")[:] = p_before
    o.state_update(2)
    assert np.allclose(o.field("prim")[inter], p_before[inter], rtol=1e-14, atol=1e-15)


def test_force_coefficients_symmetric_pressure(cloud):
    """p = f(x) on the symmetric NACA0012 wall: the y-force and pitching moment cancel
    in mirror pairs, so Cl = -tan(aoa) Cd; uniform freestream p gives zero forces."""
    o = Oracle(cloud, mach=0.63, aoa_deg=2.0).setup()
    o.compute_cl_cd_cm()
    assert o.field("Cl")[0] == 0 and o.field("Cd")[0] == 0 and o.field("Cm")[0] == 0
    p = o.field("prim")
    p[:, 3] = 1 / 1.4 + 0.1 * np.cos(3.0 * cloud.x) + 0.05 * cloud.x
    o.compute_cl_cd_cm()
    cl, cd, cm = o.field("Cl")[0], o.field("Cd")[0], o.field("Cm")[0]
    th = np.deg2rad(2.0)
    assert abs(cd) > 1e-4
    assert abs(cl + np.tan(th) * cd) < 1e-12
    assert abs(cm) < 1e-12
    cp = o.field("cp")
    assert np.allclose(cp, -(p[cloud.flag_2 > 0, 3] - 1 / 1.4) / (0.5 * 0.63 ** 2), rtol=1e-13)


def test_residue_definition(cloud):
    """fpi_solver.F90:63-70: residue 0 for t <= 2, then log10(res_new/res_old) with
    res_old frozen at iteration 2; res_new = sqrt(sum over wall+interior)/N."""
    o = Oracle(cloud, cfl=0.2, mach=0.63, aoa_deg=2.0).setup()
    res, rn = o.run(6)
    assert res[0] == 0 and res[1] == 0
    assert np.allclose(res[2:], np.log10(rn[2:] / rn[1]), rtol=1e-14)
    s = o.field("res_sqr_pt")
    assert np.isclose(np.sqrt(s[cloud.flag_1 != 2].sum()) / cloud.n, rn[-1], rtol=1e-13)


def test_nan_exit(cloud):
    o = Oracle(cloud, cfl=0.2).setup()
    o.field("prim")[5, 3] = -1.0
    res, _ = o.run(20)
    assert len(res) == 3 and np.isnan(res[-1])


def test_first_order_is_lsq_of_point_fluxes(cloud):
    """first order: q-tilde = q, so each half's derivative is the plain LSQ of
    G(q_k) - G(q_i) with the point's (t, n) frame fluxes; checked for one interior
    point against a numpy solve of the 2x2 normal equations."""
    o = Oracle(cloud, cfl=0.2, mach=0.63, aoa_deg=2.0, first_order=1).setup()
    o.run(5)
    o.func_delta()
    o.eval_q_variables()
    o.eval_q_derivatives()
    o.cal_flux_residual()
    i = int(np.where(cloud.flag_1 == 1)[0][37])
    prim = o.field("prim")
    nx, ny = o.field("nx")[i], o.field("ny")[i]
    total = np.zeros(4)
    for half, kind, ydir in (("xpos", 0, False), ("xneg", 1, False), ("ypos", 2, True), ("yneg", 3, True)):
        lst = o.ifield(half + "_conn")[i, :o.ifield(half + "_nbhs")[i]]
        def G(j):
            r, u1, u2, p = prim[j]
            return flux(kind, nx, ny, u1, u2, r, p)
        dx = cloud.x[lst] - cloud.x[i]
        dy = cloud.y[lst] - cloud.y[i]
        ds, dn = dx * ny - dy * nx, dx * nx + dy * ny
        A = np.array([[ds @ ds, ds @ dn], [ds @ dn, dn @ dn]])
        dG = np.array([G(j) - G(i) for j in lst])
        sol = np.linalg.solve(A, np.vstack([ds @ dG, dn @ dG]))
        total += sol[1] if ydir else sol[0]
    assert np.allclose(o.field("flux_res")[i], o.field("delta")[i] * total, rtol=1e-9, atol=1e-14)
