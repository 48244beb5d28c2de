#!/usr/bin/env python3
"""Benchmark of the LSKUM fixed-point iteration on one MI355X.

Workload (one "step" = one fpi_solver iteration over the whole cloud): NACA0012
O-cloud 640 x 250 = 160,000 points, 24-neighbour cap, second order with the
Venkatakrishnan limiter (K = 5), 3 inner q-derivative iterations, M 0.85, AoA 1 deg,
CFL 0.2, SSPRK(4,3) — synthetic cloud from mfcfd_amd.cloud, freestream start.

Both precision paths run the same W + K iterations from freestream; the timed region
is the K iterations (inputs resident in HBM, HIP graphs replayed back to back).
``value`` is the mixed path's point-iterations/s when its Cl and Cd stay within 1e-4
of the FP64 path's, otherwise the FP64 path's.  The roofline figures are for the
dominant kernel (the fused flux + state update), timed right after the timed region
on the state it left: 20 back-to-back launches between two HIP events on the
solver's stream (mf_time_kernels); the CPU baseline times the FP64 oracle (a C++ restatement of the
reference, OpenMP) on a few iterations of the same cloud.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(n_wall=640, n_radial=250, mach=0.85, aoa=1.0, cfl=0.2, vl_const=5.0,
                inner_iterations=3, seed=0)


def flux_update_bytes_per_point(state_bytes: int, mean_nbhs: float) -> float:
    """Algorithmic HBM bytes of one fused flux+update launch, per point.

    Reads of the point's own records (q 4, dq 8, qm 8, prim 4, prim_old 4 words of
    state; eps 1 state word; xy, nrm 4 doubles; half-stencil solve coefficients 12
    doubles, or 8 floats in the mixed path; delta 1 double; nbhs 4 B; conn 4 B per
    slot) and writes (prim, q: 8 state words).  Neighbour gathers are served from cache
    and not counted.
    """
    w = state_bytes
    coef = 12 * 8 if w == 8 else 8 * 4
    reads = (4 + 8 + 8 + 4 + 4 + 1) * w + (4 + 1) * 8 + coef + 4 + 4 * mean_nbhs
    writes = 8 * w
    return reads + writes


def iteration_bytes_per_point(state_bytes: int, mean_nbhs: float, inner: int, stages: int = 4) -> float:
    """Algorithmic bytes per point-iteration, B = S*[(1+I)(2g+20w+4N) + (5g+29w+5N)]
    + (2g+9w+4N) with g = 8 (geometry) and w the state word size."""
    g, w, N = 8.0, float(state_bytes), mean_nbhs
    return stages * ((1 + inner) * (2 * g + 20 * w + 4 * N) + (5 * g + 29 * w + 5 * N)) + (2 * g + 9 * w + 4 * N)


def run_gpu(cloud, precision: str, warmup: int, steps: int, wl: dict = WORKLOAD):
    from mfcfd_amd.solver import LskumSolver
    p = {k: wl[k] for k in ("mach", "aoa", "cfl", "vl_const", "inner_iterations")}
    s = LskumSolver(cloud, precision=precision, **p)
    print(f"[bench] {precision}: warmup {warmup}", file=sys.stderr, flush=True)
    s.q_lskum(warmup, first_it=1)
    tune = s.autotune()
    print(f"[bench] {precision}: lanes per point {tune}", file=sys.stderr, flush=True)
    s.sync()
    t0 = time.perf_counter()
    s.run_async(steps, first_it=warmup + 1)
    s.sync()
    dt = time.perf_counter() - t0
    # each stencil kernel's steady launch time on the state the timed region left
    # (back-to-back launches between HIP events on the solver's stream)
    steady = s.time_kernels(20)
    res, rn, cd = s.q_lskum(1, first_it=warmup + steps + 1)
    cl, cdv, cm, _ = s.compute_cl_cd_cm()
    prof = s.profile_kernels(3)
    lay = s.layout()
    out = dict(seconds=dt, ms_per_step=1e3 * dt / steps,
               point_iters_per_s=cloud.n * steps / dt, cl=float(cl[0]), cd=float(cdv[0]),
               residue=float(res[-1]) if len(res) else None, kernels=prof, steady_us=steady,
               layout=lay, tuning=tune)
    s.close()
    print(f"[bench] {precision}: {out['ms_per_step']:.3f} ms/iter "
          f"{out['point_iters_per_s']:.3e} pt-it/s", file=sys.stderr, flush=True)
    return out


def cpu_baseline(cloud, iters: int):
    from oracle.oracle import Oracle, num_threads
    o = Oracle(cloud, cfl=WORKLOAD["cfl"], mach=WORKLOAD["mach"], aoa_deg=WORKLOAD["aoa"],
               vl_const=WORKLOAD["vl_const"], inner_iterations=WORKLOAD["inner_iterations"]).setup()
    print(f"[bench] cpu baseline: {iters} oracle iterations", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    o.run(iters)
    dt = time.perf_counter() - t0
    return dict(value=cloud.n * iters / dt, unit="point-iters/s", cores=num_threads(),
                kind="port",
                sample=f"{iters} FP64 iterations of the same {cloud.n}-point cloud by the C++ "
                       f"oracle (restatement of src_mpi_primal), OpenMP over points, "
                       f"{dt:.1f} s")


# 2 inner iterations: with 3 the iteration on this cloud drifts upwards after ~750
# iterations (the oracle does the same); with 1 or 2 it converges over 4000
ADJ_WORKLOAD = dict(n_wall=320, n_radial=125, mach=0.8, aoa=1.25, cfl=0.2, vl_const=5.0,
                    inner_iterations=2, seed=0, iterations=2000, snaps=32)


def run_adjoint(iters: int, snaps: int):
    """Cd adjoint of the NACA0012 320x125 cloud (40,000 points), second order, 2 inner
    iterations, M 0.8, AoA 1.25: `iters` primal iterations with `snaps` binomial
    checkpoints, then the reverse sweep.  Cost is reported against a plain primal run
    of the same length."""
    from mfcfd_amd.cloud import naca_ogrid
    from mfcfd_amd.solver import LskumSolver
    w = ADJ_WORKLOAD
    cloud = naca_ogrid(w["n_wall"], w["n_radial"], seed=w["seed"])
    p = {k: w[k] for k in ("mach", "aoa", "cfl", "vl_const", "inner_iterations")}
    with LskumSolver(cloud, **p) as s:
        s.q_lskum(2)
    with LskumSolver(cloud, **p) as s:
        tune = s.autotune()
        s.sync()
        t0 = time.perf_counter()
        s.run_async(iters)
        s.sync()
        primal_s = time.perf_counter() - t0
    out = {"workload": f"NACA0012 320x125 ({cloud.n} points), 2nd order, 2 inner, M 0.8, AoA 1.25, "
                       f"J = Cd, {iters} iterations", "primal_s": primal_s,
           "tuning": {k: tune[k] for k in ("lpp_lsq", "lpp_inner", "lpp_flux")}}
    # the configured binomial schedule; every iterate kept (iters x 32 B per point of
    # HBM: the reverse re-runs each iteration once, taped); every iteration taped in
    # the primal sweep (about 1.3 kB per point per iteration: nothing re-run)
    from mfcfd_amd._lib import MF_ADJ_TAPE_ALL
    for key, ns in (("binomial", snaps), ("store_all", iters), ("tape_all", MF_ADJ_TAPE_ALL)):
        with LskumSolver(cloud, **p) as s:
            s.tuning = out["tuning"]
            print(f"[bench] adjoint: {iters} iterations, {key} ({ns})", file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            J, dx, dy, st = s.adjoint(iters, snaps=ns)
            adj_s = time.perf_counter() - t0
        out[key] = {"snapshots": ns, "J": J, "adjoint_total_s": adj_s,
                    "adjoint_over_primal": adj_s / primal_s, "stats": st,
                    "grad_norm": float(np.sqrt((dx * dx + dy * dy).sum()))}
    return out


def _boundary_normals(cloud):
    """Normals as the device computes them (point_normals.F90:8-105): wall and outer
    points from their left/right neighbours, interior points (0, 1)."""
    nx, ny = np.zeros(cloud.n), np.ones(cloud.n)
    b = cloud.flag_1 != 1
    l, r = cloud.left[b], cloud.right[b]
    x, y = cloud.x, cloud.y
    vx = 0.5 * ((y[b] - y[l]) + (y[r] - y[b]))
    vy = 0.5 * ((x[b] - x[l]) + (x[r] - x[b]))
    det = np.sqrt(vx * vx + vy * vy)
    nx[b], ny[b] = -vx / det, vy / det
    return nx, ny


def run_adjoint_fd(iters: int, n_points: int = 16, h: float = 1e-6, save: str | None = None):
    """dCd/dx, dCd/dy of the adjoint workload against central differences (step h
    chord) at seeded wall points: each difference re-runs the FP64 primal from
    freestream on the perturbed cloud (normals and half-stencil splits recomputed by
    mf_create).  A point is skipped when a perturbation flips a half-stencil split
    (delta_s or delta_n changing sign), checked before any run.  Over a long
    trajectory a perturbation of 1e-6 can also switch limiter or argmin branches (J is
    then not differentiable along the step); such points are kept and show up as the
    tail of the error distribution."""
    import dataclasses
    from mfcfd_amd.cloud import naca_ogrid, split_stencils
    from mfcfd_amd.solver import LskumSolver
    w = ADJ_WORKLOAD
    cloud = naca_ogrid(w["n_wall"], w["n_radial"], seed=w["seed"])
    p = {k: w[k] for k in ("mach", "aoa", "cfl", "vl_const", "inner_iterations")}
    with LskumSolver(cloud, **p) as s:
        J, dJdx, dJdy, _ = s.adjoint(iters, snaps=iters)
    if save:
        np.savez(save, J=J, dJdx=dJdx, dJdy=dJdy, iters=iters)
    masks0 = split_stencils(cloud, *_boundary_normals(cloud))

    def perturbed(j, comp, shift):
        c = dataclasses.replace(cloud, x=cloud.x.copy(), y=cloud.y.copy())
        (c.x if comp == 0 else c.y)[j] += shift
        return c

    def flips(c):
        return not np.array_equal(split_stencils(c, *_boundary_normals(c)), masks0)

    def J_of(c):
        with LskumSolver(c, **p) as s:
            _, _, cd = s.q_lskum(iters)
        return float(cd[-1])

    rng = np.random.default_rng(16)
    wall = np.where(cloud.flag_1 == 0)[0]
    scale = float(np.abs(np.concatenate([dJdx[wall], dJdy[wall]])).max())
    rows, flipped = [], 0
    for j in rng.permutation(wall):
        cs = {(comp, sg): perturbed(j, comp, sg * h) for comp in (0, 1) for sg in (1, -1)}
        if any(flips(c) for c in cs.values()):
            flipped += 1
            continue
        ent = {"point": int(j)}
        for comp, grad in ((0, dJdx), (1, dJdy)):
            fd = (J_of(cs[(comp, 1)]) - J_of(cs[(comp, -1)])) / (2 * h)
            ent["xy"[comp]] = {"adjoint": float(grad[j]), "fd": fd,
                               "rel_err": abs(grad[j] - fd) / max(abs(fd), 1e-3 * scale)}
        rows.append(ent)
        print(f"[bench] fd point {j}: x {ent['x']['rel_err']:.2e} y {ent['y']['rel_err']:.2e}",
              file=sys.stderr, flush=True)
        if len(rows) == n_points:
            break
    errs = np.array([e[c]["rel_err"] for e in rows for c in "xy"])
    return {"workload": f"adjoint workload, {iters} iterations, central differences h = {h}",
            "J": J, "points_used": len(rows), "skipped_split_flip": flipped,
            "rel_err_median": float(np.median(errs)), "rel_err_worst": float(errs.max()),
            "within_1e-6": int((errs <= 1e-6).sum()), "within_1e-4": int((errs <= 1e-4).sum()),
            "of": int(errs.size), "gradient_scale": scale, "points": rows}


# 2,560,000 points; the configuration fixes no inner count and no limiter constant:
# 0 inner iterations (readcase.F90's default) and K = 5 as the 160k workload
LARGE_WORKLOAD = dict(n_wall=2560, n_radial=1000, mach=1.2, aoa=0.0, cfl=0.2, vl_const=5.0,
                      inner_iterations=0, seed=0, warmup=100, steps=500)


def run_large():
    """NACA0012 2560x1000 (2.56M points), second order, M 1.2 (bow shock), AoA 0:
    100 untimed + 500 timed iterations on each precision path."""
    from mfcfd_amd.cloud import naca_ogrid
    w = LARGE_WORKLOAD
    t0 = time.perf_counter()
    cloud = naca_ogrid(w["n_wall"], w["n_radial"], seed=w["seed"], verbose=True)
    print(f"[bench] large cloud {cloud.n} points in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    nb = float(cloud.nbhs.mean())
    out = {"workload": f"NACA0012 O-cloud 2560x1000 ({cloud.n} points), 2nd order, K=5, 0 inner, "
                       f"M 1.2, AoA 0, CFL 0.2, {w['warmup']} + {w['steps']} timed iterations",
           "mean_nbhs": round(nb, 3)}
    for prec in ("fp64", "mixed"):
        r = run_gpu(cloud, prec, w["warmup"], w["steps"], wl=w)
        sb = 8 if prec == "fp64" else 4
        fu = r["steady_us"]["flux_update"] * 1e-3
        per_launch = flux_update_bytes_per_point(sb, nb) * cloud.n
        out[prec] = {"point_iters_per_s": r["point_iters_per_s"], "ms_per_step": r["ms_per_step"],
                     "cl": r["cl"], "cd": r["cd"], "residue": r["residue"], "tuning": r["tuning"],
                     "flux_update_ms": fu, "flux_update_hbm_gbs": per_launch / (fu * 1e-3) / 1e9,
                     "iteration_hbm_gbs": iteration_bytes_per_point(sb, nb, w["inner_iterations"])
                     * r["point_iters_per_s"] / 1e9,
                     "kernel_ms": {k: v[0] for k, v in r["kernels"].items()}}
    out["mixed_over_fp64"] = out["mixed"]["point_iters_per_s"] / out["fp64"]["point_iters_per_s"]
    out["dCl"] = abs(out["mixed"]["cl"] - out["fp64"]["cl"])
    out["dCd"] = abs(out["mixed"]["cd"] - out["fp64"]["cd"])
    return out


# 64 cases: 8 NACA 00tt sections (t ~ U[0.08, 0.15]) on clouds of 20k-120k points
# (wall 100-480, radial 200-250), each flown at 8 conditions M ~ U[0.5, 0.85],
# AoA ~ U[-2, 4] (seeded); second order, K = 5, 2 inner iterations, CFL 0.2
BATCH_WORKLOAD = dict(seed=5, n_airfoils=8, n_cases=64, cfl=0.2, vl_const=5.0, inner_iterations=2,
                      iterations=1000, warmup=20)


def batch_cases():
    from mfcfd_amd.cloud import naca_ogrid
    w = BATCH_WORKLOAD
    rng = np.random.default_rng(w["seed"])
    th = rng.uniform(0.08, 0.15, w["n_airfoils"])
    nw = rng.integers(100, 481, w["n_airfoils"])
    nr = rng.integers(200, 251, w["n_airfoils"])
    mach = rng.uniform(0.5, 0.85, w["n_cases"])
    aoa = rng.uniform(-2.0, 4.0, w["n_cases"])
    clouds = []
    for a in range(w["n_airfoils"]):
        t0 = time.perf_counter()
        clouds.append(naca_ogrid(int(nw[a]), int(nr[a]), thickness=float(th[a]), seed=a))
        print(f"[bench] batch cloud {a}: NACA00{round(100 * th[a]):02d} {nw[a]}x{nr[a]} "
              f"({clouds[-1].n} points) {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    cases = [(clouds[k % w["n_airfoils"]], float(mach[k]), float(aoa[k])) for k in range(w["n_cases"])]
    return cases, th


def run_batch(iters: int, adjoint: bool = True):
    """All 64 cases in one context: every graph-replayed iteration is one launch per
    kernel over all cases.  Mixed and FP64 paths from freestream for `iters`
    iterations; then the FP64 Cd adjoint of the batch over the same iterations."""
    from mfcfd_amd.solver import LskumBatch
    w = BATCH_WORKLOAD
    cases, th = batch_cases()
    p = {k: w[k] for k in ("cfl", "vl_const", "inner_iterations")}
    npts = sum(c.n for c, _, _ in cases)
    out = {"workload": f"{w['n_cases']} cases on {w['n_airfoils']} NACA 00tt clouds (thickness "
                       f"{th.min():.3f}-{th.max():.3f}, {min(c.n for c, _, _ in cases)}-"
                       f"{max(c.n for c, _, _ in cases)} points), M 0.5-0.85, AoA -2..4, 2nd order, "
                       f"K=5, 2 inner, CFL 0.2, {iters} iterations", "points": npts}
    cd_of = {}
    for prec in ("mixed", "fp64"):
        with LskumBatch(cases, precision=prec, **p) as b:
            b.run(w["warmup"])
            tune = b.autotune()
            b.sync()
            t0 = time.perf_counter()
            res, _, _ = b.run(iters, first_it=w["warmup"] + 1)
            dt = time.perf_counter() - t0
            _, cd, _, _ = b.compute_cl_cd_cm()
            cd_of[prec] = cd[:, 0]
            out[prec] = {"point_iters_per_s": npts * iters / dt, "ms_per_step": 1e3 * dt / iters,
                         "tuning": {k: tune[k] for k in ("lpp_lsq", "lpp_inner", "lpp_flux")},
                         "final_residue_max": float(np.nanmax(res[-1]))}
            print(f"[bench] batch {prec}: {out[prec]['ms_per_step']:.2f} ms/iter "
                  f"{out[prec]['point_iters_per_s']:.3e} pt-it/s", file=sys.stderr, flush=True)
    dcd = np.abs(cd_of["mixed"] - cd_of["fp64"])
    out["mixed_over_fp64"] = out["mixed"]["point_iters_per_s"] / out["fp64"]["point_iters_per_s"]
    out["per_case_dCd_mixed_vs_fp64"] = {"max": float(dcd.max()), "median": float(np.median(dcd))}
    out["cd_fp64"] = [round(float(v), 8) for v in cd_of["fp64"]]
    if adjoint:
        # every iterate kept when it fits comfortably in HBM (32 B per point per iterate)
        snaps = int(min(iters + w["warmup"], max(32, 150e9 // (npts * 32))))
        os.environ["MFCFD_PROGRESS"] = "1"
        with LskumBatch(cases, **p) as b:
            b.run(0)
            t0 = time.perf_counter()
            J, dx, dy, st = b.adjoint(iters + w["warmup"], snaps=snaps)
            adt = time.perf_counter() - t0
        fp64_iter_s = out["fp64"]["ms_per_step"] * 1e-3
        out["adjoint"] = {"iterations": iters + w["warmup"], "snapshots": snaps, "J_sum_cd": J,
                          "seconds": adt, "over_primal": adt / (fp64_iter_s * (iters + w["warmup"])),
                          "stats": st, "grad_norm": float(np.sqrt((dx * dx + dy * dy).sum()))}
    return out


COPY_PROBE_BYTES = 256 << 20
# the launch shapes the autotuner keeps for this workload (16 | L: staged walk, L lanes)
PROBE_SHAPES = {"fp64": {"lpp_lsq": 16 | 4, "lpp_inner": 16 | 4, "lpp_flux": 16 | 4},
                "mixed": {"lpp_lsq": 16 | 4, "lpp_inner": 16 | 4, "lpp_flux": 16 | 4}}


def probe():
    import ctypes as C
    from mfcfd_amd import _lib
    from mfcfd_amd.cloud import naca_ogrid
    from mfcfd_amd.solver import LskumSolver
    ms = C.c_double()
    _lib.check(_lib.lib().mf_copy_probe(COPY_PROBE_BYTES, 1, C.byref(ms)), "mf_copy_probe")
    cloud = naca_ogrid(WORKLOAD["n_wall"], WORKLOAD["n_radial"], seed=WORKLOAD["seed"])
    p = {k: WORKLOAD[k] for k in ("mach", "aoa", "cfl", "vl_const", "inner_iterations")}
    for prec in ("fp64", "mixed"):
        with LskumSolver(cloud, precision=prec, **p) as s:
            s.tuning = PROBE_SHAPES[prec]
            s.q_lskum(2)
    print(json.dumps({"probe": "ok", "copy_bytes": COPY_PROBE_BYTES, "copy_ms": ms.value,
                      "points": cloud.n, "mean_nbhs": float(cloud.nbhs.mean())}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-iters", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-iters", type=int, default=None,
                    help="iterations of the adjoint cost measurement (0: skip; default 200, "
                         "or the workload's full 2000 with --adjoint)")
    ap.add_argument("--adjoint", action="store_true",
                    help="only the adjoint measurement")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 2.56M-point measurement reported beside the headline")
    ap.add_argument("--large", action="store_true", help="only the 2.56M-point measurement")
    ap.add_argument("--batch", action="store_true", help="only the 64-case batched measurement")
    ap.add_argument("--adjoint-fd", action="store_true",
                    help="only the adjoint's central-difference check at 16 wall points")
    ap.add_argument("--fd-step", type=float, default=1e-6, help="step of that check (chord)")
    ap.add_argument("--save-gradient", default=None, help="write the adjoint's gradient (npz)")
    ap.add_argument("--batch-iters", type=int, default=None, help="iterations of the batched run")
    ap.add_argument("--probe", action="store_true",
                    help="counter-calibration pass: one 256 MiB copy kernel, then 2 FP64 and "
                         "2 mixed iterations of the workload at the autotuner's usual shapes "
                         "(for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the LSKUM path here is single-GPU "
              "(multi-GPU halo exchange is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.probe:
        probe()
        return
    if args.adjoint:
        iters = ADJ_WORKLOAD["iterations"] if args.adjoint_iters is None else args.adjoint_iters
        print(json.dumps(run_adjoint(iters, ADJ_WORKLOAD["snaps"])))
        return
    if args.large:
        print(json.dumps(run_large()))
        return
    if args.batch:
        print(json.dumps(run_batch(args.batch_iters or BATCH_WORKLOAD["iterations"])))
        return
    if args.adjoint_fd:
        iters = ADJ_WORKLOAD["iterations"] if args.adjoint_iters is None else args.adjoint_iters
        print(json.dumps(run_adjoint_fd(iters, h=args.fd_step, save=args.save_gradient)))
        return
    if args.adjoint_iters is None:
        args.adjoint_iters = 200

    from mfcfd_amd.cloud import naca_ogrid
    t0 = time.perf_counter()
    cloud = naca_ogrid(WORKLOAD["n_wall"], WORKLOAD["n_radial"], seed=WORKLOAD["seed"])
    print(f"[bench] cloud {cloud.n} points in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    mean_nbhs = float(cloud.nbhs.mean())

    fp64 = run_gpu(cloud, "fp64", args.warmup, args.steps)
    mixed = run_gpu(cloud, "mixed", args.warmup, args.steps)
    dcl = abs(mixed["cl"] - fp64["cl"])
    dcd = abs(mixed["cd"] - fp64["cd"])
    use_mixed = dcl <= 1e-4 and dcd <= 1e-4
    best = mixed if use_mixed else fp64
    sb = 4 if use_mixed else 8

    fu_ms = best["steady_us"]["flux_update"] * 1e-3
    per_launch = flux_update_bytes_per_point(sb, mean_nbhs) * cloud.n
    achieved = per_launch / (fu_ms * 1e-3) / 1e9
    traffic = valu_busy = l2_hit = traffic_kernel = None
    tpath = os.path.join(ROOT, "profiles", "traffic_flux_update.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        key = "fp32" if use_mixed else "fp64"
        traffic = t.get(key)
        valu_busy = t.get(f"{key}_valu_busy")
        l2_hit = t.get(f"{key}_l2_hit")
        traffic_kernel = t.get("kernels", {}).get(key)

    it_bytes = iteration_bytes_per_point(sb, mean_nbhs, WORKLOAD["inner_iterations"])
    line = {
        "metric": "point-iters/s + HBM GB/s, FP32-mixed vs FP64 at dCl<=1e-4; adjoint/primal cost",
        "value": best["point_iters_per_s"],
        "unit": "point-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if use_mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "NACA0012 O-cloud 640x250 (160,000 points), 2nd order, "
                               "Venkatakrishnan K=5, 3 inner iterations, M 0.85, AoA 1, CFL 0.2",
                   "points": cloud.n, "mean_nbhs": round(mean_nbhs, 3),
                   "precision": "mixed" if use_mixed else "fp64", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_flux_update", "kernel_ms": fu_ms,
                     "algorithmic_bytes_per_launch": per_launch,
                     # counters of the same kernel (profiles/): the VALU, not HBM, binds it
                     "valu_busy_pct": valu_busy, "l2_hit": l2_hit,
                     # the launch shape timed here (lanes per point, 16 | L: staged) and the
                     # kernel instance the counters were taken on
                     "shape": best["tuning"]["lpp_flux"], "counters_kernel": traffic_kernel},
        "iteration_hbm_gbs": it_bytes * best["point_iters_per_s"] / 1e9,
        "fp64_point_iters_per_s": fp64["point_iters_per_s"],
        "mixed_point_iters_per_s": mixed["point_iters_per_s"],
        "mixed_over_fp64": mixed["point_iters_per_s"] / fp64["point_iters_per_s"],
        "mixed_vs_fp64": {"dCl": dcl, "dCd": dcd, "cl_fp64": fp64["cl"], "cd_fp64": fp64["cd"]},
        "kernel_ms": {"fp64": {k: v[0] for k, v in fp64["kernels"].items()},
                      "mixed": {k: v[0] for k, v in mixed["kernels"].items()}},
        "steady_us": {"fp64": fp64["steady_us"], "mixed": mixed["steady_us"]},
    }
    if args.adjoint_iters > 0:
        line["adjoint"] = run_adjoint(args.adjoint_iters, ADJ_WORKLOAD["snaps"])
    if not args.no_large:
        line["large"] = run_large()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cloud, args.cpu_iters)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
