#!/usr/bin/env python3
"""Benchmark of the LSKUM fixed-point iteration on one MI355X.

Headline workload (one "step" = one fpi_solver iteration over the whole cloud):
NACA0012 O-cloud 2560 x 1000 = 2,560,000 points, 24-neighbour cap, second order with
the Venkatakrishnan limiter (K = 5), 0 inner iterations, M 1.2 (bow shock), AoA 0,
CFL 0.2, SSPRK(4,3); synthetic cloud from mfcfd_amd.cloud, freestream start
(mfcfd_amd/workloads.py).

Each precision path runs W warm-up iterations, autotunes its launch shapes, and then
times exactly K iterations (inputs resident in HBM, HIP graphs replayed back to back,
a stream synchronisation on both sides).  It then continues to the workload's own
schedule, 100 untimed + 500 timed iterations, reported beside the headline, and the
mixed path's Cl and Cd are compared with the FP64 path's after those 600 iterations.
``value`` is the mixed path's point-iterations/s over the K timed iterations when
|dCl|, |dCd| <= 1e-4 there, otherwise the FP64 path's.

The roofline entry is for the dominant kernel (the fused flux residual + state
update): its algorithmic bytes per point follow the scope's counting rule for the
flux + update phase, 5g + 29w + 5N (g = 8-byte geometry words, w = state word, N =
mean neighbours), and its duration is the mean of 20 back-to-back launches between two
HIP events on the solver's stream right after the timed region (mf_time_kernels).  The
kernel is bound by its VALU issue rate, not by HBM; the VALU counters of the same
kernel (profiles/) are reported beside it.

Extra keys: the 160,000-point workload over its full 5000 iterations, the 64-case
batch (primal), the Cd adjoint of the 40,000-point workload over 2000 iterations, and
the CPU baselines (the reference's own compiled routines on one core, and the C++
oracle over OpenMP), each on a bounded sample of the headline workload.

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

from mfcfd_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0


def flux_phase_bytes_per_point(state_bytes: int, mean_nbhs: float) -> float:
    """Algorithmic HBM bytes per point of one flux + update phase: 5g + 29w + 5N.

    Counting rule: each per-point array of the phase once (geometry 5 words of g = 8
    bytes; 29 state words of w bytes: q 4, dq 8, qm 8, prim 4, prim_old 4, eps 1 read
    and prim, q 8 written, less the 8 counted as the update's outputs; connectivity
    4 bytes per slot plus a 1-byte direction mask, N slots), neighbour gathers served
    from cache."""
    return 5 * 8 + 29 * state_bytes + 5 * mean_nbhs


def iteration_bytes_per_point(state_bytes: int, mean_nbhs: float, inner: int, stages: int = 4) -> float:
    """Algorithmic bytes per point-iteration, B = S*[(1+I)(2g+20w+4N) + (5g+29w+5N)]
    + (2g+9w+4N) with g = 8 (geometry) and w the state word size."""
    g, w, N = 8.0, float(state_bytes), mean_nbhs
    return stages * ((1 + inner) * (2 * g + 20 * w + 4 * N) + (5 * g + 29 * w + 5 * N)) + (2 * g + 9 * w + 4 * N)


def _timed(s, iters: int, first_it: int) -> float:
    s.sync()
    t0 = time.perf_counter()
    s.run_async(iters, first_it=first_it)
    s.sync()
    return time.perf_counter() - t0


def state_errors(prim: np.ndarray, prim_ref: np.ndarray, cp: np.ndarray, cp_ref: np.ndarray) -> dict:
    """The mixed path's state against the FP64 path's at the same iteration: per
    primitive component (rho, u1, u2, p) the largest and the root-mean-square point
    difference, each over that component's largest magnitude; cp's largest difference
    over the wall points."""
    names = ("rho", "u1", "u2", "p")
    d = np.abs(prim - prim_ref)
    scale = np.abs(prim_ref).max(axis=0)
    out = {"prim_max": {n: float(d[:, r].max() / scale[r]) for r, n in enumerate(names)},
           "prim_rms": {n: float(np.sqrt(np.mean(d[:, r] ** 2)) / scale[r]) for r, n in enumerate(names)}}
    out["prim_max_all"] = max(out["prim_max"].values())
    out["cp_max_abs"] = float(np.abs(cp - cp_ref).max()) if len(cp) else None
    out["cp_range"] = float(cp_ref.max() - cp_ref.min()) if len(cp) else None
    return out


def run_headline(cloud, params: dict, precision: str, warmup: int, steps: int, wl: dict):
    """W warm-up, autotune, K timed iterations (the headline), the kernels' steady
    launch times, then the workload's 100 + 500 schedule; Cl, Cd at the end."""
    from mfcfd_amd.solver import LskumSolver
    s = LskumSolver(cloud, precision=precision, **params)
    print(f"[bench] {precision}: warmup {warmup}", file=sys.stderr, flush=True)
    s.q_lskum(warmup, first_it=1)
    tune = s.autotune()
    print(f"[bench] {precision}: lanes per point {tune}", file=sys.stderr, flush=True)
    dt = _timed(s, steps, warmup + 1)
    # each stencil kernel's steady launch time on the state the timed region left
    # (back-to-back launches between HIP events on the solver's stream)
    steady = s.time_kernels(20)
    done = warmup + steps
    if done < wl["warmup"]:
        s.run_async(wl["warmup"] - done, first_it=done + 1)
        done = wl["warmup"]
    dt_sched = _timed(s, wl["steps"], done + 1)
    done += wl["steps"]
    res, _, _ = s.q_lskum(1, first_it=done + 1)
    done += 1
    cl, cdv, _, cp = s.compute_cl_cd_cm()
    final_prim = s.prim
    prof = s.profile_kernels(3)
    out = dict(final_prim=final_prim, cp=cp, ms_per_step=1e3 * dt / steps, point_iters_per_s=cloud.n * steps / dt,
               schedule={"iterations": f"{wl['warmup']} untimed + {wl['steps']} timed",
                         "ms_per_step": 1e3 * dt_sched / wl["steps"],
                         "point_iters_per_s": cloud.n * wl["steps"] / dt_sched},
               iterations_at_gate=done, cl=float(cl[0]), cd=float(cdv[0]),
               residue=float(res[-1]) if len(res) else None,
               kernel_ms={k: v[0] for k, v in prof.items()}, steady_us=steady, tuning=tune)
    s.close()
    print(f"[bench] {precision}: {out['ms_per_step']:.3f} ms/iter {out['point_iters_per_s']:.3e} pt-it/s "
          f"(schedule {out['schedule']['ms_per_step']:.3f} ms/iter)", file=sys.stderr, flush=True)
    return out


def run_c2():
    """The 160,000-point workload over its full 5000 iterations on both paths, timed
    whole; Cl, Cd agreement at the end."""
    from mfcfd_amd.solver import LskumSolver
    cloud, p, n_it = W.cloud_of(W.C2), W.C2["params"], W.C2["iterations"]
    out = {"workload": W.DESCRIPTION["c2"] + f", {n_it} iterations from freestream",
           "points": cloud.n, "mean_nbhs": round(float(cloud.nbhs.mean()), 3)}
    state = {}
    for prec in ("fp64", "mixed"):
        with LskumSolver(cloud, precision=prec, **p) as s:
            tune = s.autotune()
            dt = _timed(s, n_it - 1, 1)
            res, _, _ = s.q_lskum(1, first_it=n_it)
            cl, cd, _, cp = s.compute_cl_cd_cm()
            state[prec] = (s.prim, cp)
            steady = s.time_kernels(20)
        out[prec] = {"ms_per_step": 1e3 * dt / (n_it - 1), "point_iters_per_s": cloud.n * (n_it - 1) / dt,
                     "cl": float(cl[0]), "cd": float(cd[0]), "final_residue": float(res[-1]),
                     "steady_us": steady, "tuning": {k: tune[k] for k in ("lpp_lsq", "lpp_inner", "lpp_flux")}}
        print(f"[bench] c2 {prec}: {out[prec]['ms_per_step']:.3f} ms/iter", file=sys.stderr, flush=True)
    out["mixed_over_fp64"] = out["mixed"]["point_iters_per_s"] / out["fp64"]["point_iters_per_s"]
    out["dCl"] = abs(out["mixed"]["cl"] - out["fp64"]["cl"])
    out["dCd"] = abs(out["mixed"]["cd"] - out["fp64"]["cd"])
    out["mixed_state_vs_fp64"] = state_errors(state["mixed"][0], state["fp64"][0], state["mixed"][1],
                                              state["fp64"][1])
    return out


def cpu_baseline_port(cloud, params: dict, iters: int):
    """The C++ oracle (a restatement of src_mpi_primal, OpenMP over points) on
    `iters` FP64 iterations of the headline cloud."""
    from oracle.oracle import Oracle, num_threads
    o = Oracle(cloud, **W.oracle_params(params)).setup()
    print(f"[bench] cpu baseline (port): {iters} oracle iterations", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    o.run(iters)
    dt = time.perf_counter() - t0
    return dict(value=cloud.n * iters / dt, unit="point-iters/s", cores=num_threads(), kind="port",
                sample=f"{iters} FP64 iterations of the same {cloud.n}-point cloud by the C++ oracle "
                       f"(restatement of src_mpi_primal), OpenMP over points, {dt:.1f} s")


def cpu_baseline_reference(cloud, params: dict, prim: np.ndarray):
    """The reference's own compiled routines (src_mpi_primal built unmodified by
    oracle/Makefile into oracle/_ref; single-threaded, as one MPI rank of the reference
    runs) on the headline cloud from a developed state: one SSPRK stage's
    eval_q_variables, eval_q_derivatives and cal_flux_residual (0 inner iterations),
    each timed.  An iteration is 4 such stages plus func_delta and state_update, which
    include PETSc headers and are not built; the point-iterations/s below count the
    4 stages' routines only, so they favour the CPU slightly."""
    from oracle.oracle import Oracle, RefRoutines
    if not RefRoutines.available():
        return None
    o = Oracle(cloud, **W.oracle_params(params)).setup()
    ref = RefRoutines(cloud, vl_const=params["vl_const"])
    ref.set_split(o)
    ref.call("compute_normals")
    del o
    ref.set("prim", prim)
    o_delta = np.full(cloud.n, 0.0)
    times = {}
    for r in ("eval_q_variables", "eval_q_derivatives", "cal_flux_residual"):
        if r == "cal_flux_residual":
            # a time step per point (its value does not change the work)
            o_delta[:] = 1e-3
            ref.set("delta", o_delta)
        print(f"[bench] cpu baseline (reference): {r}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        ref.call(r)
        times[r] = time.perf_counter() - t0
    stage = sum(times.values())
    return dict(value=cloud.n / (4 * stage), unit="point-iters/s", cores=1, kind="reference",
                sample=f"one SSPRK stage of the reference's compiled routines (src_mpi_primal "
                       f"q_variables, flux_residual and the flux modules, amdflang -O2, 1 core) on the "
                       f"same {cloud.n}-point cloud from the FP64 state after the GPU run: "
                       + ", ".join(f"{k} {v:.2f} s" for k, v in times.items())
                       + "; point-iterations/s = N / (4 x stage); func_delta and state_update "
                         "(PETSc-bound, not built) excluded",
                seconds=times)


def cpu_adjoint_reference(cloud, params: dict, prim: np.ndarray):
    """The reference's Tapenade reverse routines (src_mpi_adjoint's *_b modules built
    unmodified by oracle/Makefile into oracle/_tap; single-threaded) on the adjoint
    workload's cloud from its final state: one reversed SSPRK stage's
    cal_flux_residual_b, inner iterations (eval_update_innerloop_b + eval_q_inner_loop_b
    each), eval_q_derivatives_b and eval_q_variables_b, each timed.  A reversed
    iteration is 4 such stages plus state_update_b and func_delta_b, which include PETSc
    headers and are not built."""
    from oracle.oracle import Oracle, TapRoutines
    if not TapRoutines.available():
        return None
    o = Oracle(cloud, **W.oracle_params(params)).setup()
    t = TapRoutines(cloud, vl_const=params["vl_const"])
    t.set_split(o)
    del o
    t.call("compute_normals")
    t.set("prim", prim.reshape(-1))
    t.call("eval_q_variables")
    t.call("eval_q_derivatives")
    t.set("delta", np.full(cloud.n, 1e-3))
    rng = np.random.default_rng(7)
    n, inner = cloud.n, params["inner_iterations"]
    times = {}

    def timed(name, routines, seeds):
        t.zero_adjoints()
        for f, size in seeds:
            t.set_adj(f, rng.standard_normal(size * n))
        print(f"[bench] cpu adjoint (Tapenade routines): {name}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for r in routines:
            t.call(r)
        times[name] = time.perf_counter() - t0

    timed("cal_flux_residual_b", ["cal_flux_residual_b"], [("flux_res", 4)])
    timed("inner_iteration_b", ["eval_update_innerloop_b", "eval_q_inner_loop_b"], [("dq", 8)])
    timed("eval_q_derivatives_b", ["eval_q_derivatives_b"], [("dq", 8), ("qm", 8)])
    timed("eval_q_variables_b", ["eval_q_variables_b"], [("q", 4)])
    stage = times["cal_flux_residual_b"] + inner * times["inner_iteration_b"] + \
        times["eval_q_derivatives_b"] + times["eval_q_variables_b"]
    return dict(seconds_per_reversed_iteration=4 * stage, cores=1, kind="reference",
                sample=f"one reversed SSPRK stage of the reference's Tapenade routines (src_mpi_adjoint "
                       f"*_b modules, amdflang -O2, 1 core) on the {n}-point cloud from the adjoint's "
                       f"final state, {inner} inner iterations: "
                       + ", ".join(f"{k} {v:.3f} s" for k, v in times.items())
                       + "; a reversed iteration = 4 stages (state_update_b and func_delta_b, "
                         "PETSc-bound, not built, excluded)",
                seconds=times)


def run_adjoint(iters: int, snaps: int, cpu: bool = True):
    """Cd adjoint of the 40,000-point workload: `iters` primal iterations with `snaps`
    binomial checkpoints, then the reverse sweep; also every iterate kept and every
    iteration taped.  Cost is reported against a plain primal run of the same
    length."""
    from mfcfd_amd.solver import LskumSolver
    cloud, p = W.cloud_of(W.C3), W.C3["params"]
    with LskumSolver(cloud, **p) as s:
        s.q_lskum(2)
    with LskumSolver(cloud, **p) as s:
        tune = s.autotune()
        primal_s = _timed(s, iters, 1)
    out = {"workload": W.DESCRIPTION["c3"] + f", {iters} iterations", "primal_s": primal_s,
           "tuning": {k: tune[k] for k in ("lpp_lsq", "lpp_inner", "lpp_flux")}}
    # the configured binomial schedule; every iterate kept (iters x 32 B per point of
    # HBM: the reverse re-runs each iteration once, taped); every iteration taped in
    # the primal sweep (about 1.3 kB per point per iteration: nothing re-run)
    from mfcfd_amd._lib import MF_ADJ_TAPE_ALL
    grads = {}
    for key, ns in (("binomial", snaps), ("store_all", iters), ("tape_all", MF_ADJ_TAPE_ALL)):
        with LskumSolver(cloud, **p) as s:
            s.tuning = out["tuning"]
            print(f"[bench] adjoint: {iters} iterations, {key} ({ns})", file=sys.stderr, flush=True)
            t0 = time.perf_counter()
            J, dx, dy, st = s.adjoint(iters, snaps=ns)
            adj_s = time.perf_counter() - t0
            final_prim = s.prim
        grads[key] = (J, np.concatenate([dx, dy]))
        out[key] = {"snapshots": ns, "J": J, "adjoint_total_s": adj_s,
                    "adjoint_over_primal": adj_s / primal_s, "stats": st,
                    "grad_norm": float(np.sqrt((dx * dx + dy * dy).sum()))}
    # the mixed-precision adjoint (FP32 primal and taped re-runs, FP64 reverse), every
    # iteration taped, against the FP64 one; its cost against the mixed primal too
    with LskumSolver(cloud, precision="mixed", **p) as s:
        tm = s.autotune()
        mixed_primal_s = _timed(s, iters, 1)
    with LskumSolver(cloud, precision="mixed", **p) as s:
        s.tuning = {k: tm[k] for k in ("lpp_lsq", "lpp_inner", "lpp_flux")}
        print(f"[bench] adjoint: {iters} iterations, mixed, tape_all", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        Jm, mx, my, st = s.adjoint(iters, snaps=MF_ADJ_TAPE_ALL)
        adj_s = time.perf_counter() - t0
    J0, g0 = grads["tape_all"]
    gm = np.concatenate([mx, my])
    v = np.random.default_rng(123).standard_normal(g0.size)
    out["mixed_tape_all"] = {"J": Jm, "adjoint_total_s": adj_s, "mixed_primal_s": mixed_primal_s,
                             "adjoint_over_fp64_primal": adj_s / primal_s,
                             "adjoint_over_mixed_primal": adj_s / mixed_primal_s, "stats": st,
                             "vs_fp64": {"J_rel": abs(Jm - J0) / abs(J0),
                                         "grad_max_rel": float(np.abs(gm - g0).max() / np.abs(g0).max()),
                                         "grad_l2_rel": float(np.linalg.norm(gm - g0) / np.linalg.norm(g0)),
                                         "directional_rel": abs(float((gm - g0) @ v)) / abs(float(g0 @ v))}}
    # the device tangent along one seeded direction, against the FP64 adjoint's
    vx, vy = v[:cloud.n], v[cloud.n:]
    with LskumSolver(cloud, **p) as s:
        t0 = time.perf_counter()
        Jt, dJt = s.tangent(iters, vx, vy)
        tan_s = time.perf_counter() - t0
    out["tangent"] = {"J": Jt, "dJ": dJt, "seconds": tan_s, "over_primal": tan_s / primal_s,
                      "adjoint_directional": float(g0 @ v),
                      "rel_vs_adjoint": abs(dJt - float(g0 @ v)) / abs(float(g0 @ v))}
    if cpu:
        ref = cpu_adjoint_reference(cloud, p, final_prim)
        if ref is not None:
            # the GPU's reverse sweep per iteration, every iteration taped (no re-runs)
            gpu_rev = 1e-3 * out["tape_all"]["stats"]["adjoint_ms"] / iters
            ref["gpu_seconds_per_reversed_iteration"] = gpu_rev
            ref["gpu_over_cpu"] = ref["seconds_per_reversed_iteration"] / gpu_rev
            out["cpu_adjoint"] = ref
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
    from mfcfd_amd.cloud import split_stencils
    from mfcfd_amd.solver import LskumSolver
    cloud, p = W.cloud_of(W.C3), W.C3["params"]
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


def run_batch(iters: int, adjoint: bool = True, warmup: int = 20):
    """All 64 cases in one context: every graph-replayed iteration is one launch per
    kernel over all cases.  Mixed and FP64 paths from freestream: `warmup` iterations,
    then `iters` timed; then (optionally) the Cd adjoint of the batch over the same
    iterations in both precisions, the mixed one's per-case gradients against FP64."""
    from mfcfd_amd.solver import LskumBatch
    cases, th = W.c5_cases()
    p = W.C5["params"]
    npts = sum(c.n for c, _, _ in cases)
    out = {"workload": W.DESCRIPTION["c5"] + f" (thickness {th.min():.3f}-{th.max():.3f}, "
                       f"{min(c.n for c, _, _ in cases)}-{max(c.n for c, _, _ in cases)} points), "
                       f"{warmup} + {iters} iterations", "points": npts}
    cd_of = {}
    for prec in ("mixed", "fp64"):
        with LskumBatch(cases, precision=prec, **p) as b:
            b.run(warmup)
            tune = b.autotune()
            b.sync()
            t0 = time.perf_counter()
            res, _, _ = b.run(iters, first_it=warmup + 1)
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
        # the Cd adjoint of the batch in both precisions (every iterate kept when it fits
        # comfortably in HBM: 32 B per point per iterate), per case against FP64
        snaps = int(min(iters + warmup, max(32, 150e9 // (npts * 32))))
        os.environ["MFCFD_PROGRESS"] = "1"
        grads = {}
        out["adjoint"] = {"iterations": iters + warmup, "snapshots": snaps}
        for prec in ("fp64", "mixed"):
            with LskumBatch(cases, precision=prec, **p) as b:
                # the forward shapes the primal run's autotuner kept
                b.tuning = out[prec]["tuning"]
                b.run(0)
                t0 = time.perf_counter()
                J, dx, dy, st = b.adjoint(iters + warmup, snaps=snaps)
                adt = time.perf_counter() - t0
                slices = [b.case_slice(k) for k in range(len(cases))]
            grads[prec] = (dx, dy)
            it_s = out[prec]["ms_per_step"] * 1e-3
            out["adjoint"][prec] = {"J_sum_cd": J, "seconds": adt,
                                    "over_primal": adt / (it_s * (iters + warmup)), "stats": st}
            print(f"[bench] batch adjoint {prec}: {adt:.1f} s", file=sys.stderr, flush=True)
        errs = []
        for sl in slices:
            g0 = np.concatenate([grads["fp64"][0][sl], grads["fp64"][1][sl]])
            g1 = np.concatenate([grads["mixed"][0][sl], grads["mixed"][1][sl]])
            errs.append(float(np.abs(g1 - g0).max() / np.abs(g0).max()))
        out["adjoint"]["per_case_dCd_dxy_mixed_vs_fp64"] = {"max": max(errs), "median": float(np.median(errs))}
    return out


COPY_PROBE_BYTES = 256 << 20


def probe(workload: str):
    """Counter-calibration pass (for rocprofv3 --pmc): one 256 MiB copy kernel, then
    2 FP64 and 2 mixed iterations of the workload at the launch shapes the autotuner
    keeps for it (autotuned first, outside the counted kernels' names)."""
    import ctypes as C
    from mfcfd_amd import _lib
    from mfcfd_amd.solver import LskumSolver
    ms = C.c_double()
    _lib.check(_lib.lib().mf_copy_probe(COPY_PROBE_BYTES, 1, C.byref(ms)), "mf_copy_probe")
    w = {"c2": W.C2, "c4": W.C4}[workload]
    cloud, p = W.cloud_of(w), w["params"]
    shapes = {}
    for prec in ("fp64", "mixed"):
        with LskumSolver(cloud, precision=prec, **p) as s:
            t = s.autotune()
            shapes[prec] = {k: t[k] for k in ("lpp_lsq", "lpp_inner", "lpp_flux")}
            s.q_lskum(2)
    print(json.dumps({"probe": "ok", "workload": workload, "copy_bytes": COPY_PROBE_BYTES,
                      "copy_ms": ms.value, "points": cloud.n, "shapes": shapes,
                      "mean_nbhs": float(cloud.nbhs.mean())}))


def run_inner_shapes():
    """Every launch shape of the mixed path's inner sweeps on the 40,000- and the
    160,000-point workloads: the staged kernel (one launch per sweep) with 1, 2, 4 lanes
    per point and the resident sweeps (every sweep of a stage in one launch, a grid
    barrier between sweeps) where their blocks fit.  Per shape: the time per sweep over
    20 back-to-back launches, and 100 iterations timed whole."""
    from mfcfd_amd._lib import MfError
    from mfcfd_amd.solver import LskumSolver
    out = {}
    for name, w in (("c3", W.C3), ("c2", W.C2)):
        cloud, p = W.cloud_of(w), w["params"]
        res = {"workload": W.DESCRIPTION[name], "inner_iterations": p["inner_iterations"]}
        for kind, flag in (("staged", 16), ("resident", 32)):
            for lanes in (1, 2, 4):
                key = f"{kind}_{lanes}"
                with LskumSolver(cloud, precision="mixed", **p) as s:
                    try:
                        s.tuning = {"lpp_lsq": 16 | 2, "lpp_flux": 16 | 4, "lpp_inner": flag | lanes}
                    except MfError:
                        res[key] = None
                        continue
                    s.q_lskum(5)
                    us = s.time_kernels(20)["inner"]
                    dt = _timed(s, 100, 6)
                    res[key] = {"us_per_sweep": us, "ms_per_iter": 1e3 * dt / 100}
                print(f"[bench] inner shapes {name} {key}: {res[key]}", file=sys.stderr, flush=True)
        out[name] = res
    return out


def _counters(key: str, prec: str):
    """Calibrated counters of the flux + update kernel of one workload (profiles/)."""
    tpath = os.path.join(ROOT, "profiles", "traffic_flux_update.json")
    if not os.path.exists(tpath):
        return {}
    with open(tpath) as f:
        return json.load(f).get(key, {}).get(prec, {})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-iters", type=int, default=4, help="oracle iterations of the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-iters", type=int, default=None,
                    help="iterations of the adjoint cost measurement (0: skip; default the "
                         "workload's 2000)")
    ap.add_argument("--adjoint", action="store_true", help="only the adjoint measurement")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline: skip the 160k-point, batch and adjoint figures")
    ap.add_argument("--batch", action="store_true", help="only the 64-case batched measurement")
    ap.add_argument("--batch-adjoint", action="store_true", help="with --batch: also the batch's Cd adjoint")
    ap.add_argument("--adjoint-fd", action="store_true",
                    help="only the adjoint's central-difference check at 16 wall points")
    ap.add_argument("--fd-step", type=float, default=1e-6, help="step of that check (chord)")
    ap.add_argument("--save-gradient", default=None, help="write the adjoint's gradient (npz)")
    ap.add_argument("--batch-iters", type=int, default=None, help="iterations of the batched run")
    ap.add_argument("--inner-shapes", action="store_true",
                    help="time every inner-sweep shape of the mixed path on the 40,000- and "
                         "160,000-point workloads and print that block only")
    ap.add_argument("--probe", choices=("c2", "c4"), default=None,
                    help="counter-calibration pass on a workload (for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the LSKUM path here is single-GPU "
              "(multi-GPU halo exchange is design-only); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.inner_shapes:
        print(json.dumps({"inner_sweep_shapes": run_inner_shapes()}))
        return
    if args.probe:
        probe(args.probe)
        return
    adj_iters = W.C3["iterations"] if args.adjoint_iters is None else args.adjoint_iters
    if args.adjoint:
        print(json.dumps(run_adjoint(adj_iters, W.C3["snaps"], cpu=not args.no_cpu)))
        return
    if args.batch:
        print(json.dumps(run_batch(args.batch_iters or W.C5["iterations"], adjoint=args.batch_adjoint)))
        return
    if args.adjoint_fd:
        print(json.dumps(run_adjoint_fd(adj_iters, h=args.fd_step, save=args.save_gradient)))
        return

    wl = W.C4
    cloud, params = W.cloud_of(wl, verbose=False), wl["params"]
    mean_nbhs = float(cloud.nbhs.mean())
    fp64 = run_headline(cloud, params, "fp64", args.warmup, args.steps, wl)
    mixed = run_headline(cloud, params, "mixed", args.warmup, args.steps, wl)
    dcl = abs(mixed["cl"] - fp64["cl"])
    dcd = abs(mixed["cd"] - fp64["cd"])
    use_mixed = dcl <= 1e-4 and dcd <= 1e-4
    best = mixed if use_mixed else fp64
    prec = "mixed" if use_mixed else "fp64"
    sb = 4 if use_mixed else 8

    fu_s = best["steady_us"]["flux_update"] * 1e-6
    per_point = flux_phase_bytes_per_point(sb, mean_nbhs)
    per_launch = per_point * cloud.n
    achieved = per_launch / fu_s / 1e9
    ctr = _counters("naca0012_2560x1000", prec)
    valu_insts = ctr.get("valu_insts")
    # wave64 VALU instructions per SIMD per cycle (1,024 SIMDs at 2.4 GHz) at this
    # launch time; the nominal rate is one per 4 cycles with packed FP32 off
    issue = valu_insts / (1024 * 2.4e9 * fu_s) if valu_insts else None
    it_bytes = iteration_bytes_per_point(sb, mean_nbhs, params["inner_iterations"])
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
        "config": {"workload": W.DESCRIPTION["c4"], "points": cloud.n, "mean_nbhs": round(mean_nbhs, 3),
                   "precision": prec, "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": ctr.get("hbm_bytes"),
                     "kernel": "k_flux_update", "kernel_us": best["steady_us"]["flux_update"],
                     "algorithmic_bytes_per_point": per_point, "algorithmic_bytes_per_launch": per_launch,
                     # the VALU binds this kernel: its issued VALU instructions (counters,
                     # profiles/) per SIMD-cycle at the launch time measured here
                     "valu": {"busy_pct": ctr.get("valu_busy"), "insts_per_launch": valu_insts,
                              "insts_per_simd_cycle": issue, "nominal_per_simd_cycle": 0.25,
                              "frac_of_nominal": issue / 0.25 if issue else None},
                     "l2_hit": ctr.get("l2_hit"), "shape": best["tuning"]["lpp_flux"],
                     "counters_kernel": ctr.get("kernel")},
        "iteration_hbm_gbs": it_bytes * best["point_iters_per_s"] / 1e9,
        "fp64_point_iters_per_s": fp64["point_iters_per_s"],
        "mixed_point_iters_per_s": mixed["point_iters_per_s"],
        "mixed_over_fp64": mixed["point_iters_per_s"] / fp64["point_iters_per_s"],
        "mixed_vs_fp64": {"iterations": mixed["iterations_at_gate"], "dCl": dcl, "dCd": dcd,
                          "cl_fp64": fp64["cl"], "cd_fp64": fp64["cd"],
                          **state_errors(mixed["final_prim"], fp64["final_prim"], mixed["cp"], fp64["cp"])},
        "schedule_100_500": {"fp64": fp64["schedule"], "mixed": mixed["schedule"]},
        "kernel_ms": {"fp64": fp64["kernel_ms"], "mixed": mixed["kernel_ms"]},
        "steady_us": {"fp64": fp64["steady_us"], "mixed": mixed["steady_us"]},
        "tuning": {"fp64": fp64["tuning"], "mixed": mixed["tuning"]},
    }
    if not args.no_extra:
        line["c2"] = run_c2()
        line["batch"] = run_batch(W.C5["iterations"], adjoint=False)
        if adj_iters > 0:
            line["adjoint"] = run_adjoint(adj_iters, W.C3["snaps"], cpu=not args.no_cpu)
    if not args.no_cpu:
        ref = cpu_baseline_reference(cloud, params, fp64["final_prim"])
        port = cpu_baseline_port(cloud, params, args.cpu_iters)
        line["cpu_baseline"] = ref if ref is not None else port
        line["cpu_baseline_port"] = port
    print(json.dumps(line))


if __name__ == "__main__":
    main()
