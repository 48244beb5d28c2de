"""Writes tests/golden/ref_routines.npz: input/output vectors of the reference's own
Fortran routines (src_mpi_primal, compiled unmodified by oracle/Makefile into
oracle/_ref/libmfcfd_ref.so).

Run in a container that has the reference sources (``make -C oracle ref`` first):
    python tests/golden/make_golden.py

Contents (all float64 / int32, no pickles):
  cloud_*          a 32 x 10 NACA0012 O-cloud (320 points, seed 11, first spacing
                   0.02, far field at 4 chords)
  split_*          the oracle's half-stencil lists (handed to the reference flux
                   routines as input: the reference's generate_connectivity needs
                   PETSc, so these are no reference output; the tests pin them
                   against a brute-force split instead)
  prim, delta      a developed state (25 oracle iterations from freestream, M 0.63,
                   AoA 2, CFL 0.2, VL_CONST 5, 1 inner iteration) and its time steps
  nx, ny           reference compute_normals
  q, dq, qm        reference eval_q_variables, eval_q_derivatives
  dq_inner         dq after reference eval_q_inner_loop + eval_update_innerloop
  flux_res         reference cal_flux_residual on (q, dq_inner, qm, delta)
  flux_in/flux_out the eight reference split/quadrant fluxes on 64 random states
  qt_in/qt_out     reference qtilde_to_primitive on 64 random q-tilde vectors
  venkat_*         reference venkat_limiter on 64 (point, q-tilde) pairs
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mfcfd_amd.cloud import naca_ogrid  # noqa: E402
from oracle.oracle import Oracle, RefRoutines  # noqa: E402

PARAMS = dict(cfl=0.2, mach=0.63, aoa_deg=2.0, vl_const=5.0, inner_iterations=1)


def main():
    if not RefRoutines.available():
        raise SystemExit("oracle/_ref/libmfcfd_ref.so missing: run `make -C oracle ref`")
    cloud = naca_ogrid(32, 10, seed=11, first_spacing=0.02, far_radius=4.0)
    o = Oracle(cloud, **PARAMS).setup()
    o.run(25)
    o.func_delta()
    prim = o.field("prim").copy()
    delta = o.field("delta").copy()

    ref = RefRoutines(cloud, vl_const=PARAMS["vl_const"])
    ref.set_split(o)
    ref.call("compute_normals")
    out = {}
    out["nx"], out["ny"] = ref.get("nx"), ref.get("ny")
    ref.set("prim", prim)
    ref.call("eval_q_variables")
    out["q"] = ref.get("q")
    ref.call("eval_q_derivatives")
    out["dq"], out["qm"] = ref.get("dq"), ref.get("qm")
    ref.call("eval_q_inner_loop")
    ref.call("eval_update_innerloop")
    out["dq_inner"] = ref.get("dq")
    ref.set("delta", delta)
    ref.call("cal_flux_residual")
    out["flux_res"] = ref.get("flux_res")

    rng = np.random.default_rng(7)
    fin = np.column_stack([
        rng.uniform(-1, 1, 64), rng.uniform(-1, 1, 64),      # normal direction (normalised)
        rng.uniform(-1.5, 1.5, 64), rng.uniform(-1.5, 1.5, 64),  # u1, u2
        rng.uniform(0.3, 2.0, 64), rng.uniform(0.2, 1.5, 64)])   # rho, pr
    nrm = np.hypot(fin[:, 0], fin[:, 1])
    fin[:, 0] /= nrm
    fin[:, 1] /= nrm
    fout = np.array([[ref.flux(k, *row) for k in range(8)] for row in fin])
    qt_in = np.column_stack([rng.uniform(-3, 1, 64), rng.uniform(-2, 2, 64),
                             rng.uniform(-2, 2, 64), -rng.uniform(0.5, 3.0, 64)])
    qt_out = np.array([ref.qtilde_to_primitive(v) for v in qt_in])
    # limiter on the reference's current q/qm; q-tilde = q(k) + perturbations of
    # varying size so both the phi = 1 branch and the limited branch are taken
    vk = rng.integers(0, cloud.n, 64)
    scale = 10.0 ** rng.uniform(-7, -1, (64, 1))
    v_qt = out["q"][vk] + scale * rng.standard_normal((64, 4))
    v_phi = np.array([ref.venkat(int(k), qt) for k, qt in zip(vk, v_qt)])

    split = {f"split_{f}": o.ifield(f).astype(np.int32).copy() for f in (
        "xpos_nbhs", "xneg_nbhs", "ypos_nbhs", "yneg_nbhs",
        "xpos_conn", "xneg_conn", "ypos_conn", "yneg_conn")}
    np.savez_compressed(
        os.path.join(HERE, "ref_routines.npz"),
        cloud_x=cloud.x, cloud_y=cloud.y, cloud_nx=cloud.nx, cloud_ny=cloud.ny,
        cloud_min_dist=cloud.min_dist, cloud_left=cloud.left, cloud_right=cloud.right,
        cloud_flag_1=cloud.flag_1, cloud_flag_2=cloud.flag_2, cloud_nbhs=cloud.nbhs,
        cloud_conn=cloud.conn, prim=prim, delta=delta, flux_in=fin, flux_out=fout,
        qt_in=qt_in, qt_out=qt_out, venkat_k=vk.astype(np.int32), venkat_qt=v_qt,
        venkat_phi=v_phi, params=np.array([PARAMS["cfl"], PARAMS["mach"], PARAMS["aoa_deg"],
                                           PARAMS["vl_const"], PARAMS["inner_iterations"]]),
        **split, **out)
    print("wrote", os.path.join(HERE, "ref_routines.npz"))


if __name__ == "__main__":
    main()
