"""Host-side mirror of the reference's hot-path interface, over the C ABI.

The reference drives its iteration through argument-less module procedures
(src_mpi_primal: ``fpi_solver``, ``func_delta``, ``eval_q_variables``,
``eval_q_derivatives``, ``eval_q_inner_loop``, ``cal_flux_residual``,
``state_update``, ``compute_cl_cd_cm``) over module state.  :class:`LskumSolver`
exposes the same procedure names on an MI355X context that owns that state; arrays
in and out use the reference's layouts and original point order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .cloud import PointCloud

_FIELDS = {"q": (L.MF_FIELD_Q, (4,)), "dq": (L.MF_FIELD_DQ, (4, 2)), "qm": (L.MF_FIELD_QM, (4, 2)),
           "flux_res": (L.MF_FIELD_FLUX_RES, (4,)), "delta": (L.MF_FIELD_DELTA, ()),
           "prim_old": (L.MF_FIELD_PRIM_OLD, (4,)), "nx": (L.MF_FIELD_NX, ()),
           "ny": (L.MF_FIELD_NY, ())}

# readcase.F90:10-125 option names -> mf_params fields
_TSCHEME = {"ssprk43": (4, 1.0), "first": (1, 2.0)}


def make_params(**kw) -> L.MfParams:
    """mf_params from reference option names (cfl, mach, aoa, vl_const, power,
    inner_iterations, tscheme, shapes, interior_points_normal_flag, format, restart,
    res_old, solution_accuracy) plus precision ('fp64'|'mixed') and reorder."""
    p = L.MfParams()
    L.lib().mf_default_params(C.byref(p))
    kw = dict(kw)
    if "tscheme" in kw:
        p.rks, p.euler = _TSCHEME[kw.pop("tscheme")]
    if "aoa_deg" in kw:
        kw["aoa"] = kw.pop("aoa_deg")
    acc = kw.pop("solution_accuracy", None)
    if acc is not None:
        kw["first_order"] = {"second": 0, "first": 1}[acc]
    prec = kw.pop("precision", "fp64")
    p.precision = {"fp64": L.MF_FP64, "mixed": L.MF_MIXED}[prec] if isinstance(prec, str) else int(prec)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown option {k}")
        setattr(p, k, type(getattr(p, k))(v))
    return p


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _geometry(cloud: PointCloud):
    """mf_geometry of a cloud in the reference's conventions (1-based indices, conn(N, W)
    column-major); returns the struct and the arrays it points into."""
    conn_f = np.where(cloud.conn >= 0, cloud.conn + 1, 0).astype(np.int32)
    keep = dict(
        x=np.ascontiguousarray(cloud.x, dtype=np.float64),
        y=np.ascontiguousarray(cloud.y, dtype=np.float64),
        nx=np.ascontiguousarray(cloud.nx, dtype=np.float64),
        ny=np.ascontiguousarray(cloud.ny, dtype=np.float64),
        min_dist=np.ascontiguousarray(cloud.min_dist, dtype=np.float64),
        left=(cloud.left + 1).astype(np.int32), right=(cloud.right + 1).astype(np.int32),
        flag_1=cloud.flag_1.astype(np.int32), flag_2=cloud.flag_2.astype(np.int32),
        nbhs=cloud.nbhs.astype(np.int32), conn=np.asfortranarray(conn_f).ravel(order="F"))
    g = L.MfGeometry(n_points=cloud.n, max_nbhs=cloud.max_nbhs,
                     **{k: v.ctypes.data for k, v in keep.items()})
    return g, keep


class LskumSolver:
    """One point cloud on the GPU, with the reference's hot-path procedures."""

    def __init__(self, cloud: PointCloud, **params):
        self.cloud = cloud
        self.params = make_params(**params)
        g, self._keep = _geometry(cloud)
        h = C.c_void_p()
        L.check(L.lib().mf_create(C.byref(g), C.byref(self.params), C.byref(h)), "mf_create")
        self._init_handle(h, cloud.n)

    def _init_handle(self, h, n):
        self.h = h
        self.n = n
        self.shapes = self.params.shapes
        self.n_shape = L.lib().mf_shape_points(self.h)

    def close(self):
        if getattr(self, "h", None):
            L.lib().mf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state
    @property
    def prim(self) -> np.ndarray:
        out = np.empty((self.n, 4))
        L.check(L.lib().mf_get_prim(self.h, _dp(out)), "mf_get_prim")
        return out

    @prim.setter
    def prim(self, value: np.ndarray) -> None:
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(self.n, 4)
        L.check(L.lib().mf_set_prim(self.h, _dp(v)), "mf_set_prim")

    def field(self, name: str) -> np.ndarray:
        fid, shp = _FIELDS[name]
        out = np.empty((self.n,) + shp)
        L.check(L.lib().mf_get_field(self.h, fid, _dp(out)), f"mf_get_field({name})")
        return out

    def set_field(self, name: str, value: np.ndarray) -> None:
        fid, shp = _FIELDS[name]
        v = np.ascontiguousarray(value, dtype=np.float64).reshape((self.n,) + shp)
        L.check(L.lib().mf_set_field(self.h, fid, _dp(v)), f"mf_set_field({name})")

    # ---- reference procedures
    def func_delta(self):
        L.check(L.lib().mf_func_delta(self.h), "mf_func_delta")

    def eval_q_variables(self):
        L.check(L.lib().mf_eval_q_variables(self.h), "mf_eval_q_variables")

    def eval_q_derivatives(self):
        L.check(L.lib().mf_eval_q_derivatives(self.h), "mf_eval_q_derivatives")

    def eval_q_inner_loop(self):
        """eval_q_inner_loop followed by eval_update_innerloop."""
        L.check(L.lib().mf_eval_q_inner_loop(self.h), "mf_eval_q_inner_loop")

    def cal_flux_residual(self):
        L.check(L.lib().mf_cal_flux_residual(self.h), "mf_cal_flux_residual")

    def state_update(self, rk: int) -> float:
        s = C.c_double()
        L.check(L.lib().mf_state_update(self.h, int(rk), C.byref(s)), "mf_state_update")
        return s.value

    def compute_cl_cd_cm(self):
        """(Cl, Cd, Cm) per shape and cp per shape point (compute_force_coeffs.F90)."""
        cl, cd, cm = (np.empty(self.shapes) for _ in range(3))
        cp = np.empty(max(self.n_shape, 1))
        L.check(L.lib().mf_compute_cl_cd_cm(self.h, _dp(cl), _dp(cd), _dp(cm), _dp(cp)),
                "mf_compute_cl_cd_cm")
        return cl, cd, cm, cp[: self.n_shape]

    def q_lskum(self, max_iters: int, first_it: int = 1):
        """q_lskum.F90:59-73: returns (residue, res_new, cd) per iteration run."""
        res, rn, cd = (np.empty(max(max_iters, 1)) for _ in range(3))
        done = C.c_int32()
        L.check(L.lib().mf_run(self.h, int(first_it), int(max_iters), _dp(res), _dp(rn), _dp(cd),
                               C.byref(done)), "mf_run")
        d = done.value
        return res[:d], rn[:d], cd[:d]

    def fpi_solver(self, t: int) -> float:
        """One iteration (fpi_solver.F90:13-83); returns its residue."""
        res, _, _ = self.q_lskum(1, first_it=t)
        return float(res[0]) if len(res) else float("nan")

    def run_async(self, iters: int, first_it: int = 1):
        L.check(L.lib().mf_run_async(self.h, int(first_it), int(iters)), "mf_run_async")

    def sync(self):
        L.check(L.lib().mf_sync(self.h), "mf_sync")

    def profile_kernels(self, iters: int) -> dict:
        ms = np.zeros(L.MF_NUM_KERNEL_KINDS)
        cnt = np.zeros(L.MF_NUM_KERNEL_KINDS, dtype=np.int32)
        L.check(L.lib().mf_profile_kernels(self.h, int(iters), _dp(ms),
                                           cnt.ctypes.data_as(C.POINTER(C.c_int32))),
                "mf_profile_kernels")
        return {k: (float(ms[j]), int(cnt[j])) for j, k in enumerate(L.KERNEL_KINDS)}

    def time_kernels(self, reps: int = 20) -> dict:
        """Mean launch time (us) of the LSQ, inner-sweep and flux+update kernels at the
        current shapes, each over `reps` back-to-back launches; the state is unchanged."""
        us = np.zeros(3)
        L.check(L.lib().mf_time_kernels(self.h, int(reps), _dp(us)), "mf_time_kernels")
        return {"lsq": float(us[0]), "inner": float(us[1]), "flux_update": float(us[2])}

    # ---- launch shapes (lanes per point of the stencil kernels)
    def autotune(self, reps: int = 10) -> dict:
        """Time every lanes-per-point shape on the current state and keep the fastest
        of each stencil kernel; returns the choice and the measured launch times."""
        t = L.MfTuning()
        L.check(L.lib().mf_autotune(self.h, int(reps), C.byref(t)), "mf_autotune")
        return {f: getattr(t, f) for f, _ in L.MfTuning._fields_}

    _SHAPES = ("lpp_lsq", "lpp_inner", "lpp_flux", "lpp_flux_rev", "lpp_lsq_rev", "lpp_inner_rev",
               "lpp_gather", "barrier_polls")

    @property
    def tuning(self) -> dict:
        """Lanes per point of each stencil kernel (the reverse ones 0 = timed by the
        first adjoint run), and the polls before a wait at the resident inner sweeps'
        grid barrier gives up."""
        t = L.MfTuning()
        L.check(L.lib().mf_get_tuning(self.h, C.byref(t)), "mf_get_tuning")
        return {k: getattr(t, k) for k in self._SHAPES}

    @tuning.setter
    def tuning(self, value: dict) -> None:
        cur = self.tuning
        cur.update(value)
        t = L.MfTuning(**{k: cur[k] for k in self._SHAPES})
        L.check(L.lib().mf_set_tuning(self.h, C.byref(t)), "mf_set_tuning")

    def layout(self) -> dict:
        li = L.MfLayout()
        L.check(L.lib().mf_layout(self.h, C.byref(li)), "mf_layout")
        return {f: getattr(li, f) for f, _ in L.MfLayout._fields_}

    # ---- adjoint (src_mpi_adjoint q_lskum_chkpts_b)
    def adjoint(self, iters: int, snaps: int = 32, first_it: int = 1):
        """J = sum of Cd after `iters` iterations from the current state, and dJ/dx,
        dJ/dy per point (the reference's pointb%x, pointb%y).  Returns
        (J, dJdx, dJdy, stats).  snaps: prim snapshots of the binomial schedule, or
        L.MF_ADJ_TAPE_ALL (-1) to tape every iteration in the primal sweep."""
        J = C.c_double()
        dx, dy = np.empty(self.n), np.empty(self.n)
        st = L.MfAdjointStats()
        L.check(L.lib().mf_adjoint(self.h, int(first_it), int(iters), int(snaps), C.byref(J),
                                   _dp(dx), _dp(dy), C.byref(st)), "mf_adjoint")
        return J.value, dx, dy, {f: getattr(st, f) for f, _ in L.MfAdjointStats._fields_}

    def tangent(self, iters: int, vx, vy, prim_tangent: bool = False):
        """Forward mode: J = sum of Cd after `iters` iterations from the current state
        and its derivative along the coordinate direction (vx, vy); with prim_tangent,
        also the final prim's tangent (N, 4).  The state is left unchanged."""
        vx = np.ascontiguousarray(vx, dtype=np.float64)
        vy = np.ascontiguousarray(vy, dtype=np.float64)
        J, dJ = C.c_double(), C.c_double()
        pd = np.empty((self.n, 4)) if prim_tangent else None
        L.check(L.lib().mf_tangent(self.h, int(iters), _dp(vx), _dp(vy), C.byref(J), C.byref(dJ),
                                   _dp(pd) if pd is not None else None), "mf_tangent")
        return (J.value, dJ.value, pd) if prim_tangent else (J.value, dJ.value)

    def adjoint_forward(self, iters: int, snaps: int = 32, first_it: int = 1) -> float:
        J = C.c_double()
        L.check(L.lib().mf_adjoint_forward(self.h, int(first_it), int(iters), int(snaps),
                                           C.byref(J)), "mf_adjoint_forward")
        return J.value

    def adjoint_backward(self):
        dx, dy = np.empty(self.n), np.empty(self.n)
        st = L.MfAdjointStats()
        L.check(L.lib().mf_adjoint_backward(self.h, _dp(dx), _dp(dy), C.byref(st)),
                "mf_adjoint_backward")
        return dx, dy, {f: getattr(st, f) for f, _ in L.MfAdjointStats._fields_}


class LskumBatch(LskumSolver):
    """Several independent flow cases in one context (mf_create_batch): every launch
    advances all of them.  ``cases`` is a list of (cloud, mach, aoa); the other options
    are shared.  Per-point arrays are the cases' points concatenated in case order."""

    def __init__(self, cases, **params):
        self.params = make_params(**params)
        self._geoms = [_geometry(c) for c, _, _ in cases]
        arr = (L.MfCase * len(cases))()
        for k, ((cloud, mach, aoa), (g, _)) in enumerate(zip(cases, self._geoms)):
            arr[k].geometry = C.pointer(g)
            arr[k].mach = float(mach)
            arr[k].aoa = float(aoa)
        self._cases_arr = arr
        h = C.c_void_p()
        L.check(L.lib().mf_create_batch(arr, len(cases), C.byref(self.params), C.byref(h)), "mf_create_batch")
        self.ncases = len(cases)
        self.clouds = [c for c, _, _ in cases]
        self._init_handle(h, sum(c.n for c in self.clouds))
        first = np.zeros(self.ncases + 1, dtype=np.int32)
        nc = C.c_int32()
        L.check(L.lib().mf_batch_cases(self.h, C.byref(nc), first.ctypes.data_as(C.POINTER(C.c_int32))),
                "mf_batch_cases")
        self.first = first

    def run(self, iters: int, first_it: int = 1):
        """(residue, res_new) of shape (iters, ncases) and cd of shape (iters, ncases, shapes)."""
        res = np.empty((max(iters, 1), self.ncases))
        rn = np.empty_like(res)
        cd = np.empty((max(iters, 1), self.ncases, self.shapes))
        L.check(L.lib().mf_batch_run(self.h, int(first_it), int(iters), _dp(res), _dp(rn), _dp(cd)),
                "mf_batch_run")
        return res[:iters], rn[:iters], cd[:iters]

    def compute_cl_cd_cm(self):
        """Cl, Cd, Cm of shape (ncases, shapes) and cp of every shape point."""
        k = self.ncases * self.shapes
        cl, cd, cm = (np.empty(k) for _ in range(3))
        cp = np.empty(max(self.n_shape, 1))
        L.check(L.lib().mf_compute_cl_cd_cm(self.h, _dp(cl), _dp(cd), _dp(cm), _dp(cp)),
                "mf_compute_cl_cd_cm")
        shp = (self.ncases, self.shapes)
        return cl.reshape(shp), cd.reshape(shp), cm.reshape(shp), cp[: self.n_shape]

    def case_slice(self, k: int) -> slice:
        return slice(int(self.first[k]), int(self.first[k + 1]))
