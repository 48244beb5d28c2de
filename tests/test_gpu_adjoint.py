"""GPU adjoint (mf_adjoint) against the oracle's dual-number tangent and central finite
differences of the FP64 oracle.

* Any direction (vx, vy): sum(dJ/dx vx + dJ/dy vy) from the GPU reverse sweep equals
  the oracle's forward-mode derivative to 1e-6 relative (the bar the reference's
  Tapenade adjoint is held to).  The tangent is itself checked against central
  differences in tests/test_oracle.py (test_dual_tangent_central_difference).
* 16 seeded wall points: dJ/dx, dJ/dy against central differences of the oracle with
  a 1e-7 step (stencil split held fixed, normals recomputed), to 1e-5 relative.
* The checkpoint schedule changes the recomputation, not the result.

The reference's Tapenade reverse routines are pinned routine by routine in
tests/test_tapenade.py; its whole-iteration pointb%x/y (state_update_b, fpi_solver_b,
PETSc) cannot be built here and stays unpinned.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = dict(cfl=0.2, mach=0.63, aoa=2.0, vl_const=5.0, inner_iterations=2)


def _gpu(cloud, **kw):
    from mfcfd_amd.solver import LskumSolver
    p = dict(P)
    p.update(kw)
    return LskumSolver(cloud, **p)


def _oparams(**kw):
    p = dict(P)
    p.update(kw)
    p["aoa_deg"] = p.pop("aoa")
    return p


@pytest.mark.parametrize("variant", ["second_inner2", "first_order", "tscheme_first", "power_m2", "power_1"])
def test_adjoint_matches_tangent(small_cloud, variant):
    """Also with distance-weighted least squares (power -2 and 1: the weights'
    derivatives in the LSQ, inner-iteration and flux reverses, q_variables_b.F90)."""
    from oracle.oracle import tangent_cd
    kw = {"second_inner2": {}, "first_order": dict(first_order=1, inner_iterations=0),
          "tscheme_first": dict(inner_iterations=1), "power_m2": dict(power=-2.0),
          "power_1": dict(power=1.0)}[variant]
    okw = dict(kw)
    gkw = dict(kw)
    if variant == "tscheme_first":
        okw.update(rks=1, euler=2.0)
        gkw["tscheme"] = "first"
    iters = 20
    print(f"[adjoint] {variant}: {small_cloud.n} points, {iters} iterations", flush=True)
    with _gpu(small_cloud, **gkw) as g:
        J, dJdx, dJdy, st = g.adjoint(iters, snaps=4)
    rng = np.random.default_rng(11)
    vx, vy = rng.standard_normal(small_cloud.n), rng.standard_normal(small_cloud.n)
    Jo, dJo = tangent_cd(small_cloud, vx, vy, iters, **_oparams(**okw))
    adj = float(dJdx @ vx + dJdy @ vy)
    print(f"J {J:.14e} vs {Jo:.14e}; directional {adj:.12e} vs tangent {dJo:.12e}; stats {st}")
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert abs(adj - dJo) <= 1e-6 * abs(dJo)
    assert st["taped_steps"] == iters
    assert np.all(np.isfinite(dJdx)) and np.all(np.isfinite(dJdy))


def test_adjoint_fd_wall_points(small_cloud):
    """16 seeded wall points where J is differentiable (the one-sided differences
    agree; the trailing edge, with its exact ties, is not): central differences with a
    1e-7 step match the adjoint to 1e-5."""
    from oracle.oracle import Oracle
    iters = 20
    with _gpu(small_cloud) as g:
        J, dJdx, dJdy, _ = g.adjoint(iters, snaps=6)
    rng = np.random.default_rng(5)
    wall = np.where(small_cloud.flag_1 == 0)[0]
    h = 1e-7
    cache = {}

    def Jof(j, comp, s):
        key = (j, comp, s)
        if key not in cache:
            o = Oracle(small_cloud, **_oparams()).setup()
            o.field("x" if comp == 0 else "y")[j] += s
            o.compute_normals()
            o.run(iters)
            o.compute_cl_cd_cm()
            cache[key] = float(o.field("Cd").sum())
        return cache[key]

    used, worst = 0, 0.0
    for j in rng.permutation(wall):
        fwd = [(Jof(j, c, h) - Jof(j, c, 0.0)) / h for c in (0, 1)]
        bwd = [(Jof(j, c, 0.0) - Jof(j, c, -h)) / h for c in (0, 1)]
        if any(abs(f - b) > 1e-4 * max(abs(f), abs(b), 1e-6) for f, b in zip(fwd, bwd)):
            continue  # a kink of the primal at this point
        for comp, grad in ((0, dJdx), (1, dJdy)):
            fd = 0.5 * (fwd[comp] + bwd[comp])
            worst = max(worst, abs(grad[j] - fd) / max(abs(fd), 1e-6))
        used += 1
        if used == 16:
            break
    print(f"worst FD relative error at {used} wall points: {worst:.3e}")
    assert used == 16
    assert worst < 1e-5


def test_adjoint_schedule_invariance(small_cloud):
    """1 snapshot (recompute from the start), 3, one per iteration, and every iteration
    taped in the primal sweep (MF_ADJ_TAPE_ALL) agree; the taped sweep leaves the same
    final state and re-runs nothing."""
    from mfcfd_amd._lib import MF_ADJ_TAPE_ALL
    out, final = [], []
    for snaps in (1, 3, 25, MF_ADJ_TAPE_ALL):
        with _gpu(small_cloud) as g:
            J, dx, dy, st = g.adjoint(12, snaps=snaps)
            out.append((J, dx, dy, st))
            final.append(g.prim)
    ref = np.concatenate([out[2][1], out[2][2]])
    for k, (J, dx, dy, st) in enumerate(out):
        assert J == out[2][0]
        g = np.concatenate([dx, dy])
        assert np.abs(g - ref).max() <= 1e-10 * np.abs(ref).max()
        assert np.array_equal(final[k], final[2])
    assert out[0][3]["forward_steps"] > out[1][3]["forward_steps"] >= out[2][3]["forward_steps"]
    assert out[3][3]["forward_steps"] == 12 and out[3][3]["taped_steps"] == 12
    # a checkpointed run after a taped one on the same context, from the same state
    with _gpu(small_cloud) as g:
        start = g.prim
        J1, _, _, _ = g.adjoint(12, snaps=MF_ADJ_TAPE_ALL)
        g.prim = start
        J2, dx2, dy2, _ = g.adjoint(12, snaps=3)
        assert J1 == J2 == out[2][0]
        assert np.abs(np.concatenate([dx2, dy2]) - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("lpp", [2, 4, 8])
def test_adjoint_reverse_shapes(small_cloud, lpp):
    """The reverse kernels split over 2/4/8 lanes per point (partial sums combined with
    shuffles; the qm argmax resolved over the lanes) give the one-lane gradient up to
    summation order, and the same J."""
    out = []
    for shape in (1, lpp):
        with _gpu(small_cloud) as g:
            g.tuning = {"lpp_flux_rev": shape, "lpp_lsq_rev": shape, "lpp_inner_rev": shape,
                        "lpp_gather": shape}
            J, dx, dy, st = g.adjoint(12, snaps=4)
            assert (st["lpp_flux_rev"], st["lpp_lsq_rev"], st["lpp_inner_rev"], st["lpp_gather"]) == \
                (shape,) * 4
            out.append((J, np.concatenate([dx, dy])))
    assert out[0][0] == out[1][0]
    assert np.abs(out[1][1] - out[0][1]).max() <= 1e-12 * np.abs(out[0][1]).max()


def test_adjoint_leaves_final_state(small_cloud):
    with _gpu(small_cloud) as a, _gpu(small_cloud) as b:
        a.adjoint(10, snaps=3)
        b.q_lskum(10)
        assert np.array_equal(a.prim, b.prim)


def test_adjoint_matches_tangent_c3():
    """The 40,000-point workload over 2000 iterations against the oracle tangent stored
    in tests/golden/c3_tangent.json (every iterate kept, so no recomputation)."""
    import json
    import os
    import sys
    from mfcfd_amd.cloud import naca_ogrid
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_c3_tangent as M
    with open(os.path.join(here, "c3_tangent.json")) as f:
        ref = json.load(f)
    c = naca_ogrid(320, 125, seed=0)
    assert M.cloud_digest(c) == ref["cloud_sha256"], "cloud generator changed: remake the fixture"
    p = dict(ref["params"])
    p["aoa"] = p.pop("aoa_deg")
    from mfcfd_amd.solver import LskumSolver
    iters = ref["iterations"]
    print(f"[adjoint] c3: {c.n} points, {iters} iterations", flush=True)
    vx, vy = M.direction(c.n)
    with LskumSolver(c, **p) as g:
        Jt, dJt = g.tangent(iters, vx, vy)
        J, dJdx, dJdy, st = g.adjoint(iters, snaps=iters)
    adj = float(dJdx @ vx + dJdy @ vy)
    print(f"J {J:.15e} vs {ref['J']:.15e}; directional {adj:.14e} vs oracle tangent {ref['dJ']:.14e}; "
          f"device tangent {dJt:.14e} (J {Jt:.15e})")
    assert abs(J - ref["J"]) <= 1e-12 * abs(ref["J"])
    assert abs(adj - ref["dJ"]) <= 1e-10 * abs(ref["dJ"])
    # the device tangent mode against the adjoint: w (J v) = (J^T w) v with w = 1
    assert abs(Jt - J) <= 1e-10 * abs(J)
    assert abs(dJt - adj) <= 1e-10 * abs(adj)


@pytest.mark.parametrize("kernel,lpp", [("flux", 2), ("flux", 4), ("flux", 8), ("inner", 1), ("inner", 2),
                                        ("inner", 4), ("lsq", 1), ("lsq", 2), ("lsq", 4)])
def test_adjoint_staged_reverse(small_cloud, kernel, lpp):
    """The staged forms of the flux reverse (32-point blocks) and of the LSQ and
    inner-iteration reverses (64-point blocks) read the taped stage state from an LDS
    halo and walk the
    same slots in the same order as the global forms: the same gradient up to the order
    of the few floating-point atomics of the sweep (func_delta's argmin, the force
    integral, the normals)."""
    key = {"flux": "lpp_flux_rev", "inner": "lpp_inner_rev", "lsq": "lpp_lsq_rev"}[kernel]
    pin = {"lpp_flux_rev": 2, "lpp_lsq_rev": 2, "lpp_inner_rev": 2, "lpp_gather": 2}
    with _gpu(small_cloud) as a, _gpu(small_cloud) as b:
        a.tuning = {**pin, key: lpp}
        b.tuning = {**pin, key: 16 | lpp}
        Ja, xa, ya, sa = a.adjoint(10, snaps=3)
        Jb, xb, yb, sb = b.adjoint(10, snaps=3)
    assert sa[key] == lpp and sb[key] == 16 | lpp
    ref = np.concatenate([xa, ya])
    assert Ja == Jb and np.abs(np.concatenate([xb, yb]) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("kernel", ["inner", "lsq", "flux"])
def test_adjoint_staged_reverse_overflow(c1_cloud, kernel):
    """On the 9,600-point cloud the largest halos exceed the LDS capacity (the 99th
    percentile of the blocks' halo sizes; at most 248 entries for the 64-point halos):
    64-point blocks of wall points read their neighbours past the capacity from global
    memory, and 32-point blocks whose halo does not fit walk the global records.  The
    staged reverses still give the global walk's gradient."""
    key = {"inner": "lpp_inner_rev", "lsq": "lpp_lsq_rev", "flux": "lpp_flux_rev"}[kernel]
    pin = {"lpp_flux_rev": 2, "lpp_lsq_rev": 2, "lpp_inner_rev": 2, "lpp_gather": 2}
    with _gpu(c1_cloud) as a, _gpu(c1_cloud) as b:
        a.tuning = {**pin, key: 2}
        b.tuning = {**pin, key: 16 | 2}
        Ja, xa, ya, _ = a.adjoint(6, snaps=2)
        Jb, xb, yb, _ = b.adjoint(6, snaps=2)
    ref = np.concatenate([xa, ya])
    assert Ja == Jb and np.abs(np.concatenate([xb, yb]) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("lpp", [1, 2, 4])
@pytest.mark.parametrize("power", [0.0, -2.0])
def test_adjoint_pull_inner_reverse(small_cloud, lpp, power):
    """The pull form of the inner-iteration reverse (MF_LPP_PULL | lanes: each point
    gathers the terms it receives as a neighbour from its owners' records and solve
    adjoints staged in LDS, with no edge buffer and no gather kernel) gives the push
    form's gradient up to summation order, with and without distance weights."""
    pin = {"lpp_flux_rev": 2, "lpp_lsq_rev": 2, "lpp_gather": 2}
    with _gpu(small_cloud, power=power) as a, _gpu(small_cloud, power=power) as b:
        a.tuning = {**pin, "lpp_inner_rev": 2}
        b.tuning = {**pin, "lpp_inner_rev": 64 | lpp}
        Ja, xa, ya, _ = a.adjoint(10, snaps=3)
        Jb, xb, yb, sb = b.adjoint(10, snaps=3)
    assert sb["lpp_inner_rev"] == 64 | lpp
    ref = np.concatenate([xa, ya])
    assert Ja == Jb and np.abs(np.concatenate([xb, yb]) - ref).max() <= 1e-12 * np.abs(ref).max()


def test_adjoint_pull_inner_reverse_overflow(c1_cloud):
    """On the 9,600-point cloud some 64-point blocks' pull halos (neighbours and owners)
    exceed the LDS capacity: those entries, as slots and as transposed entries, are read
    from global memory.  Same gradient as the push form, 3 inner iterations (so the
    pull launches hand their solve adjoints on twice)."""
    pin = {"lpp_flux_rev": 2, "lpp_lsq_rev": 2, "lpp_gather": 2}
    with _gpu(c1_cloud, inner_iterations=3) as a, _gpu(c1_cloud, inner_iterations=3) as b:
        a.tuning = {**pin, "lpp_inner_rev": 2}
        b.tuning = {**pin, "lpp_inner_rev": 64 | 4}
        Ja, xa, ya, _ = a.adjoint(6, snaps=2)
        Jb, xb, yb, _ = b.adjoint(6, snaps=2)
    ref = np.concatenate([xa, ya])
    assert Ja == Jb and np.abs(np.concatenate([xb, yb]) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("lpp", [1, 2, 4])
@pytest.mark.parametrize("power,inner", [(0.0, 2), (-2.0, 2), (0.0, 0)])
def test_adjoint_pull_lsq_reverse(small_cloud, lpp, power, inner):
    """The pull form of the LSQ reverse (MF_LPP_PULL | lanes; the qm adjoints of a
    selected neighbour delivered by atomic adds) gives the push form's gradient up to
    summation order: after the push-form inner reverse, after the pull-form one (which
    hands over the solve adjoints) and with no inner iterations."""
    pin = {"lpp_flux_rev": 2, "lpp_gather": 2}
    with _gpu(small_cloud, power=power, inner_iterations=inner) as a, \
            _gpu(small_cloud, power=power, inner_iterations=inner) as b, \
            _gpu(small_cloud, power=power, inner_iterations=inner) as c:
        a.tuning = {**pin, "lpp_lsq_rev": 2, "lpp_inner_rev": 2}
        b.tuning = {**pin, "lpp_lsq_rev": 64 | lpp, "lpp_inner_rev": 2}
        c.tuning = {**pin, "lpp_lsq_rev": 64 | lpp, "lpp_inner_rev": 64 | 4}
        Ja, xa, ya, _ = a.adjoint(10, snaps=3)
        Jb, xb, yb, sb = b.adjoint(10, snaps=3)
        Jc, xc, yc, sc = c.adjoint(10, snaps=3)
    assert sb["lpp_lsq_rev"] == 64 | lpp and sc["lpp_inner_rev"] == 64 | 4
    ref = np.concatenate([xa, ya])
    assert Ja == Jb == Jc
    assert np.abs(np.concatenate([xb, yb]) - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(np.concatenate([xc, yc]) - ref).max() <= 1e-12 * np.abs(ref).max()


def test_adjoint_pull_lsq_reverse_overflow(c1_cloud):
    """The pull-form LSQ reverse on the 9,600-point cloud, whose largest pull halos
    exceed the LDS capacity (entries read from global memory): the push form's
    gradient."""
    pin = {"lpp_flux_rev": 2, "lpp_inner_rev": 2, "lpp_gather": 2}
    with _gpu(c1_cloud) as a, _gpu(c1_cloud) as b:
        a.tuning = {**pin, "lpp_lsq_rev": 2}
        b.tuning = {**pin, "lpp_lsq_rev": 64 | 4}
        Ja, xa, ya, _ = a.adjoint(6, snaps=2)
        Jb, xb, yb, _ = b.adjoint(6, snaps=2)
    ref = np.concatenate([xa, ya])
    assert Ja == Jb and np.abs(np.concatenate([xb, yb]) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("snaps", [4, -1])
def test_mixed_adjoint_close_to_fp64(small_cloud, snaps):
    """Mixed-precision adjoint: the primal sweep and the taped re-runs run the mixed
    path's FP32 kernels, their records widened into the FP64 tape; the flux reverse runs
    in FP32 on the FP32 records (as the mixed primal evaluates the fluxes), the other
    reverse kernels in FP64, and every adjoint sum is FP64.  Against the FP64 adjoint over the same
    20 iterations: J within 1e-5 relative, the gradient within 1e-3 of its largest
    entry, and a directional derivative within 1e-3.  The binomial (4 snapshots) and
    tape-everything schedules give the same J bits (the re-runs repeat the primal's
    launches) and the same gradient up to the order of the sweep's few FP64 atomics."""
    iters = 20
    with _gpu(small_cloud) as a:
        J0, gx0, gy0, _ = a.adjoint(iters, snaps=4)
    with _gpu(small_cloud, precision="mixed") as b:
        J1, gx1, gy1, st = b.adjoint(iters, snaps=snaps)
    g0 = np.concatenate([gx0, gy0])
    g1 = np.concatenate([gx1, gy1])
    rng = np.random.default_rng(3)
    v = rng.standard_normal(g0.size)
    rel_g = float(np.abs(g1 - g0).max() / np.abs(g0).max())
    rel_d = abs(float(g1 @ v - g0 @ v)) / abs(float(g0 @ v))
    print(f"[adjoint] mixed ({snaps}): J {J1:.10e} vs {J0:.10e}; gradient rel {rel_g:.2e}; "
          f"directional rel {rel_d:.2e}; stats {st}", flush=True)
    assert abs(J1 - J0) <= 1e-5 * abs(J0)
    assert rel_g <= 1e-3 and rel_d <= 1e-3
    assert np.all(np.isfinite(g1))
    if snaps == -1:
        with _gpu(small_cloud, precision="mixed") as c:
            J2, gx2, gy2, _ = c.adjoint(iters, snaps=4)
        g2 = np.concatenate([gx2, gy2])
        assert J2 == J1 and np.abs(g2 - g1).max() <= 1e-12 * np.abs(g1).max()


@pytest.mark.parametrize("variant", ["second_inner2", "first_order", "tscheme_first", "power_m2"])
def test_device_tangent_matches_oracle(small_cloud, variant):
    """mf_tangent (device forward mode, dual numbers) against the oracle's dual-number
    tangent on 20 iterations: J and the directional derivative to 1e-10 relative; the
    final prim's tangent against the oracle's is not exposed by the oracle, so it is
    checked by linearity (twice the direction, twice the tangent)."""
    from oracle.oracle import tangent_cd
    kw = {"second_inner2": {}, "first_order": dict(first_order=1, inner_iterations=0),
          "tscheme_first": dict(inner_iterations=1), "power_m2": dict(power=-2.0)}[variant]
    okw = dict(kw)
    gkw = dict(kw)
    if variant == "tscheme_first":
        okw.update(rks=1, euler=2.0)
        gkw["tscheme"] = "first"
    iters = 20
    rng = np.random.default_rng(17)
    vx, vy = rng.standard_normal(small_cloud.n), rng.standard_normal(small_cloud.n)
    print(f"[tangent] {variant}: {small_cloud.n} points, {iters} iterations", flush=True)
    with _gpu(small_cloud, **gkw) as g:
        p0 = g.prim
        J, dJ, pd = g.tangent(iters, vx, vy, prim_tangent=True)
        J2, dJ2, pd2 = g.tangent(iters, 2 * vx, 2 * vy, prim_tangent=True)
        assert np.array_equal(g.prim, p0)  # the context's state is left as it was
    Jo, dJo = tangent_cd(small_cloud, vx, vy, iters, **_oparams(**okw))
    print(f"J {J:.14e} vs {Jo:.14e}; dJ {dJ:.12e} vs oracle {dJo:.12e}", flush=True)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert abs(dJ - dJo) <= 1e-10 * abs(dJo)
    assert J2 == J and abs(dJ2 - 2 * dJ) <= 1e-13 * abs(dJ)
    assert np.abs(pd2 - 2 * pd).max() <= 1e-13 * np.abs(pd).max() and np.abs(pd).max() > 0


def test_device_tangent_dot_adjoint(small_cloud):
    """The device tangent against the device adjoint on the same 20 iterations:
    dJ(v) = sum(dJ/dx vx + dJ/dy vy) to 1e-10, for three directions."""
    iters = 20
    with _gpu(small_cloud) as g:
        J, dJdx, dJdy, _ = g.adjoint(iters, snaps=4)
    rng = np.random.default_rng(23)
    for _ in range(3):
        vx, vy = rng.standard_normal(small_cloud.n), rng.standard_normal(small_cloud.n)
        with _gpu(small_cloud) as g:
            Jt, dJt = g.tangent(iters, vx, vy)
        adj = float(dJdx @ vx + dJdy @ vy)
        assert abs(Jt - J) <= 1e-10 * abs(J)
        assert abs(dJt - adj) <= 1e-10 * abs(adj), (dJt, adj)
