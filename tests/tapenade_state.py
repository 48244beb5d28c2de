"""Helpers of test_tapenade.py: the cloud and state tests/golden/tap_routines.npz was made on."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def golden_cloud():
    """The cloud and options tests/golden/tap_routines.npz was made on."""
    from mfcfd_amd.cloud import PointCloud
    g = np.load(os.path.join(HERE, "golden", "tap_routines.npz"), allow_pickle=False)
    c = PointCloud(**{k: g["cloud_" + k] for k in ("x", "y", "nx", "ny", "min_dist", "left", "right",
                                                   "flag_1", "flag_2", "nbhs", "conn")})
    cfl, mach, aoa, vl, inner = g["params"]
    return c, dict(cfl=cfl, mach=mach, aoa_deg=aoa, vl_const=vl, inner_iterations=int(inner))


def golden_state(tap):
    c, params = golden_cloud()
    n = c.n
    st = {"q": tap["st_q"].reshape(n, 4), "qm": tap["st_qm"].reshape(n, 4, 2)}
    return st, params["vl_const"], c.min_dist


def golden_dual():
    from oracle.oracle import DualRoutines
    c, params = golden_cloud()
    return DualRoutines(c, **params)
