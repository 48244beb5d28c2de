"""Writes tests/golden/c3_tangent.json: the oracle's forward-mode (dual-number) derivative
of J = Cd after 2000 iterations of the 40,000-point adjoint workload, in one seeded
direction of the point coordinates.

The workload: NACA0012 320 x 125 O-cloud (seed 0), second order, Venkatakrishnan K = 5,
2 inner iterations, M 0.8, AoA 1.25, CFL 0.2, from freestream.  The direction: vx, vy
standard normal per point from numpy's default_rng(123), in the cloud's point order.

The GPU adjoint's sum(dJ/dx vx + dJ/dy vy) over the same run is checked against dJ
(tests/test_gpu_adjoint.py).  The oracle run takes about 18 minutes on 8 cores:
    python tests/golden/make_c3_tangent.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

ITERS = 2000
PARAMS = dict(cfl=0.2, mach=0.8, aoa_deg=1.25, vl_const=5.0, inner_iterations=2)


def cloud_digest(cloud) -> str:
    """sha256 of the cloud's coordinates and connectivity (the fixture holds only for
    the cloud it was made on)."""
    h = hashlib.sha256()
    for a in (cloud.x.astype(np.float64), cloud.y.astype(np.float64),
              cloud.nbhs.astype(np.int32), cloud.conn.astype(np.int32)):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def direction(n: int):
    rng = np.random.default_rng(123)
    return rng.standard_normal(n), rng.standard_normal(n)


def main():
    from mfcfd_amd.cloud import naca_ogrid
    from oracle.oracle import tangent_cd
    c = naca_ogrid(320, 125, seed=0)
    vx, vy = direction(c.n)
    t0 = time.time()
    J, dJ = tangent_cd(c, vx, vy, ITERS, **PARAMS)
    out = {"cloud": "naca_ogrid(320, 125, seed=0)", "points": int(c.n),
           "cloud_sha256": cloud_digest(c), "iterations": ITERS, "params": PARAMS,
           "direction": "default_rng(123): vx, vy = standard_normal(n), standard_normal(n)",
           "J": J, "dJ": dJ, "oracle_seconds": round(time.time() - t0, 1)}
    with open(os.path.join(HERE, "c3_tangent.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
