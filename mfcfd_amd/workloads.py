"""The five workloads the path is measured and tested on, defined once.

Every cloud is synthetic (mfcfd_amd.cloud: NACA 4-digit O-clouds, seeded; the
reference ships none).  Each function returns the cloud(s) and the run options in the
keyword form of mfcfd_amd.solver (``aoa`` in degrees; the oracle's keyword is
``aoa_deg``).

* c1: NACA0012 160 x 60 = 9,600 points, first order, M 0.63, AoA 2, CFL 0.2, 1000
  iterations (FP64).
* c2: NACA0012 640 x 250 = 160,000 points, second order, Venkatakrishnan K = 5,
  3 inner iterations, M 0.85, AoA 1, CFL 0.2, 5000 iterations (FP64 and mixed).
* c3: NACA0012 320 x 125 = 40,000 points, second order, K = 5, M 0.8, AoA 1.25,
  J = Cd after 2000 iterations, 32 binomial snapshots.  2 inner iterations: with 3 the
  iteration on this cloud drifts upwards after ~750 iterations (the oracle does the
  same); with 1 or 2 it converges.
* c4: NACA0012 2560 x 1000 = 2,560,000 points, second order, K = 5, M 1.2 (bow
  shock), AoA 0, 100 + 500 iterations.  The configuration fixes no inner count:
  0 (readcase.F90's default).
* c5: 64 cases on 8 NACA 00tt sections (thickness ~ U[0.08, 0.15], wall 100-480 x
  radial 200-250 points, 20k-120k points each), each section flown at 8 conditions
  M ~ U[0.5, 0.85], AoA ~ U[-2, 4] (seeded); second order, K = 5, 2 inner
  iterations, CFL 0.2, 1000 iterations.
"""
from __future__ import annotations

import functools
import sys
import time

import numpy as np

C1 = dict(n_wall=160, n_radial=60, seed=0, iterations=1000,
          params=dict(mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=0, first_order=1))
C2 = dict(n_wall=640, n_radial=250, seed=0, iterations=5000,
          params=dict(mach=0.85, aoa=1.0, cfl=0.2, vl_const=5.0, inner_iterations=3))
C3 = dict(n_wall=320, n_radial=125, seed=0, iterations=2000, snaps=32,
          params=dict(mach=0.8, aoa=1.25, cfl=0.2, vl_const=5.0, inner_iterations=2))
C4 = dict(n_wall=2560, n_radial=1000, seed=0, warmup=100, steps=500,
          params=dict(mach=1.2, aoa=0.0, cfl=0.2, vl_const=5.0, inner_iterations=0))
C5 = dict(seed=5, n_airfoils=8, n_cases=64, iterations=1000,
          params=dict(cfl=0.2, vl_const=5.0, inner_iterations=2))

DESCRIPTION = {
    "c1": "NACA0012 O-cloud 160x60 (9,600 points), 1st order, M 0.63, AoA 2, CFL 0.2",
    "c2": "NACA0012 O-cloud 640x250 (160,000 points), 2nd order, Venkatakrishnan K=5, "
          "3 inner iterations, M 0.85, AoA 1, CFL 0.2",
    "c3": "NACA0012 O-cloud 320x125 (40,000 points), 2nd order, K=5, 2 inner iterations, "
          "M 0.8, AoA 1.25, J = Cd",
    "c4": "NACA0012 O-cloud 2560x1000 (2,560,000 points), 2nd order, Venkatakrishnan K=5, "
          "0 inner iterations, M 1.2, AoA 0, CFL 0.2",
    "c5": "64 cases on 8 NACA 00tt O-clouds (20k-120k points), M 0.5-0.85, AoA -2..4, "
          "2nd order, K=5, 2 inner iterations, CFL 0.2",
}


def oracle_params(params: dict) -> dict:
    """The same options in the oracle's keywords."""
    p = dict(params)
    p["aoa_deg"] = p.pop("aoa")
    return p


@functools.lru_cache(maxsize=None)
def _ogrid(n_wall: int, n_radial: int, seed: int, thickness: float = 0.12, verbose: bool = False):
    from mfcfd_amd.cloud import naca_ogrid
    t0 = time.perf_counter()
    c = naca_ogrid(n_wall, n_radial, thickness=thickness, seed=seed, verbose=verbose)
    print(f"[workload] cloud {n_wall}x{n_radial} t={thickness:.3f} ({c.n} points) in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return c


def cloud_of(w: dict, verbose: bool = False):
    return _ogrid(w["n_wall"], w["n_radial"], w["seed"], verbose=verbose)


def c5_cases():
    """The 64 (cloud, mach, aoa) cases and the 8 sections' thicknesses."""
    w = C5
    rng = np.random.default_rng(w["seed"])
    th = rng.uniform(0.08, 0.15, w["n_airfoils"])
    nw = rng.integers(100, 481, w["n_airfoils"])
    nr = rng.integers(200, 251, w["n_airfoils"])
    mach = rng.uniform(0.5, 0.85, w["n_cases"])
    aoa = rng.uniform(-2.0, 4.0, w["n_cases"])
    clouds = [_ogrid(int(nw[a]), int(nr[a]), a, float(th[a])) for a in range(w["n_airfoils"])]
    cases = [(clouds[k % w["n_airfoils"]], float(mach[k]), float(aoa[k])) for k in range(w["n_cases"])]
    return cases, th
