"""ctypes front-ends of the test oracles (TEST INFRASTRUCTURE ONLY).

Only the test suite, the smoke check and the benchmark's CPU-baseline leg import this
module, and only to check or time the reference algorithm; the shipped library never
touches it.

* :class:`Oracle` drives ``oracle/liboracle.so``, the C++ restatement of
  src_mpi_primal's iteration (``lskum_oracle.cpp``).
* :class:`RefRoutines` drives ``oracle/_ref/libmfcfd_ref.so``, the reference's own
  Fortran routines compiled unmodified with amdflang (built only where the reference
  sources are present).
* :class:`TapRoutines` drives ``oracle/_tap/libmfcfd_tap.so``, the reference's
  Tapenade-generated reverse routines compiled the same way.
* :class:`DualRoutines` runs single routines of the oracle in dual numbers (the
  tangents a reverse routine is checked against).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libmfcfd_ref.so")


class OrcParams(C.Structure):
    _fields_ = [("cfl", C.c_double), ("vl_const", C.c_double), ("power", C.c_double),
                ("mach", C.c_double), ("aoa_deg", C.c_double), ("euler", C.c_double),
                ("rks", C.c_int32), ("inner_iterations", C.c_int32), ("shapes", C.c_int32),
                ("interior_points_normal_flag", C.c_int32), ("format", C.c_int32),
                ("restart", C.c_int32), ("restart_res_old", C.c_double),
                ("first_order", C.c_int32)]


def build(force: bool = False) -> None:
    """Compile liboracle.so (g++) and, when the reference sources exist, _ref."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])
    if os.path.isdir("/root/reference/src_mpi_primal") and (force or not os.path.exists(REF_LIB_PATH)):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])
    if os.path.isdir("/root/reference/src_mpi_adjoint") and (force or not os.path.exists(TAP_LIB_PATH)):
        subprocess.check_call(["make", "-s", "-C", HERE, "tap"])


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        lib.orc_create.restype = vp
        lib.orc_create.argtypes = [C.c_int32, C.c_int32, dp, dp, ip, ip, ip, ip, dp, ip, ip,
                                   dp, dp, C.POINTER(OrcParams)]
        for f in ("orc_destroy", "orc_initial_conditions", "orc_compute_normals",
                  "orc_generate_connectivity", "orc_func_delta", "orc_eval_q_variables",
                  "orc_eval_q_derivatives", "orc_eval_q_inner_loop",
                  "orc_eval_update_innerloop", "orc_cal_flux_residual",
                  "orc_compute_cl_cd_cm"):
            getattr(lib, f).argtypes = [vp]
            getattr(lib, f).restype = None
        lib.orc_state_update.argtypes = [vp, C.c_int32]
        lib.orc_fpi_solver.argtypes = [vp, C.c_int32]
        lib.orc_run.argtypes = [vp, C.c_int32, C.c_int32, dp, dp]
        lib.orc_run.restype = C.c_int32
        lib.orc_field.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
        lib.orc_field.restype = dp
        lib.orc_ifield.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
        lib.orc_ifield.restype = ip
        lib.orc_scalar.argtypes = [vp, C.c_char_p]
        lib.orc_scalar.restype = C.c_double
        lib.orc_flux.argtypes = [C.c_int32] + [C.c_double] * 6 + [dp]
        lib.orc_qtilde_to_primitive.argtypes = [dp, dp]
        lib.orc_num_threads.restype = C.c_int32
        lib.orc_dual_create.restype = vp
        lib.orc_dual_create.argtypes = lib.orc_create.argtypes
        lib.orc_dual_seed.argtypes = [vp, dp, dp]
        lib.orc_dual_setup.argtypes = [vp]
        lib.orc_dual_run.argtypes = [vp, C.c_int32, C.c_int32]
        lib.orc_dual_cd.argtypes = [vp, dp, dp]
        lib.orc_dual_destroy.argtypes = [vp]
        lib.orc_dual_set.argtypes = [vp, C.c_char_p, dp, dp]
        lib.orc_dual_set.restype = C.c_int64
        lib.orc_dual_get.argtypes = [vp, C.c_char_p, dp, dp]
        lib.orc_dual_get.restype = C.c_int64
        lib.orc_dual_call.argtypes = [vp, C.c_char_p]
        lib.orc_dual_call.restype = C.c_int32
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def default_params(**kw) -> dict:
    """readcase.F90:10-125 defaults (ssprk43, venkat 150, power 0, one shape)."""
    p = dict(cfl=0.0, vl_const=150.0, power=0.0, mach=0.63, aoa_deg=2.0, euler=1.0, rks=4,
             inner_iterations=0, shapes=1, interior_points_normal_flag=0, format=1,
             restart=0, restart_res_old=0.0, first_order=0)
    p.update(kw)
    return p


class Oracle:
    """FP64 restatement of the reference iteration on one cloud (original ordering)."""

    FIELDS_4 = ("prim", "prim_old", "q", "flux_res")
    FIELDS_8 = ("dq", "qm", "temp")

    def __init__(self, cloud, **params):
        lib = _load()
        self.cloud = cloud
        self.params = default_params(**params)
        prm = OrcParams(**self.params)
        n, w = cloud.n, cloud.max_nbhs
        self._keep = [np.ascontiguousarray(a) for a in (
            cloud.x.astype(np.float64), cloud.y.astype(np.float64),
            cloud.left.astype(np.int32), cloud.right.astype(np.int32),
            cloud.flag_1.astype(np.int32), cloud.flag_2.astype(np.int32),
            cloud.min_dist.astype(np.float64), cloud.nbhs.astype(np.int32),
            np.where(cloud.conn >= 0, cloud.conn, 0).astype(np.int32),
            cloud.nx.astype(np.float64), cloud.ny.astype(np.float64))]
        x, y, l, r, f1, f2, md, nb, cn, nx, ny = self._keep
        self.h = lib.orc_create(n, w, _dp(x), _dp(y), _ip(l), _ip(r), _ip(f1), _ip(f2),
                                _dp(md), _ip(nb), _ip(cn), _dp(nx), _dp(ny), C.byref(prm))
        self.n, self.w = n, w

    def __del__(self):
        if getattr(self, "h", None):
            _load().orc_destroy(self.h)
            self.h = None

    # reference routine names (no-argument subroutines over module state)
    def setup(self):
        """initial_conditions + compute_normals + generate_connectivity (q_lskum.F90:33-34)."""
        lib = _load()
        lib.orc_initial_conditions(self.h)
        lib.orc_compute_normals(self.h)
        lib.orc_generate_connectivity(self.h)
        return self

    def __getattr__(self, name):
        if name in ("func_delta", "eval_q_variables", "eval_q_derivatives", "eval_q_inner_loop",
                    "eval_update_innerloop", "cal_flux_residual", "compute_cl_cd_cm",
                    "compute_normals", "generate_connectivity", "initial_conditions"):
            f = getattr(_load(), "orc_" + name)
            return lambda: f(self.h)
        raise AttributeError(name)

    def state_update(self, rk: int):
        _load().orc_state_update(self.h, rk)

    def fpi_solver(self, t: int):
        _load().orc_fpi_solver(self.h, t)

    def run(self, iters: int, first_it: int = 1):
        """q_lskum loop (q_lskum.F90:59-73): returns (residue, res_new) histories."""
        res = np.zeros(iters)
        rn = np.zeros(iters)
        done = _load().orc_run(self.h, first_it, iters, _dp(res), _dp(rn))
        return res[:done], rn[:done]

    def field(self, name: str) -> np.ndarray:
        """A live view of one oracle array (reshaped to (N,4) or (N,4,2))."""
        ln = C.c_int64()
        ptr = _load().orc_field(self.h, name.encode(), C.byref(ln))
        if not ptr:
            raise KeyError(name)
        a = np.ctypeslib.as_array(ptr, shape=(ln.value,))
        if name in self.FIELDS_4:
            return a.reshape(self.n, 4)
        if name in self.FIELDS_8:
            return a.reshape(self.n, 4, 2)
        return a

    def ifield(self, name: str) -> np.ndarray:
        ln = C.c_int64()
        ptr = _load().orc_ifield(self.h, name.encode(), C.byref(ln))
        if not ptr:
            raise KeyError(name)
        a = np.ctypeslib.as_array(ptr, shape=(ln.value,))
        if name.endswith("_conn"):
            return a.reshape(self.n, self.w)
        return a

    def scalar(self, name: str) -> float:
        return _load().orc_scalar(self.h, name.encode())


def tangent_cd(cloud, vx, vy, iters: int, first_it: int = 1, **params):
    """Forward-mode derivative of J = sum of Cd after `iters` iterations from
    freestream, in the direction (vx, vy) of the point coordinates.  Returns (J, dJ)."""
    lib = _load()
    prm = OrcParams(**default_params(**params))
    n, w = cloud.n, cloud.max_nbhs
    keep = [np.ascontiguousarray(a) for a in (
        cloud.x.astype(np.float64), cloud.y.astype(np.float64),
        cloud.left.astype(np.int32), cloud.right.astype(np.int32),
        cloud.flag_1.astype(np.int32), cloud.flag_2.astype(np.int32),
        cloud.min_dist.astype(np.float64), cloud.nbhs.astype(np.int32),
        np.where(cloud.conn >= 0, cloud.conn, 0).astype(np.int32),
        cloud.nx.astype(np.float64), cloud.ny.astype(np.float64))]
    x, y, l, r, f1, f2, md, nb, cn, nx, ny = keep
    h = lib.orc_dual_create(n, w, _dp(x), _dp(y), _ip(l), _ip(r), _ip(f1), _ip(f2), _dp(md),
                            _ip(nb), _ip(cn), _dp(nx), _dp(ny), C.byref(prm))
    try:
        vx = np.ascontiguousarray(vx, dtype=np.float64)
        vy = np.ascontiguousarray(vy, dtype=np.float64)
        lib.orc_dual_seed(h, _dp(vx), _dp(vy))
        lib.orc_dual_setup(h)
        lib.orc_dual_run(h, first_it, iters)
        J, dJ = C.c_double(), C.c_double()
        lib.orc_dual_cd(h, C.byref(J), C.byref(dJ))
        return J.value, dJ.value
    finally:
        lib.orc_dual_destroy(h)


def num_threads() -> int:
    """OpenMP threads the oracle's point loops use."""
    return int(_load().orc_num_threads())


def flux(kind: int, nx, ny, u1, u2, rho, pr) -> np.ndarray:
    """One kinetic split flux (0 Gxp 1 Gxn 2 Gyp 3 Gyn 4 GxI 5 GxII 6 GxIII 7 GxIV)."""
    G = np.zeros(4)
    _load().orc_flux(kind, nx, ny, u1, u2, rho, pr, _dp(G))
    return G


def qtilde_to_primitive(qt) -> np.ndarray:
    """Returns (rho, u1, u2, pr)."""
    qt = np.ascontiguousarray(qt, dtype=np.float64)
    out = np.zeros(4)
    _load().orc_qtilde_to_primitive(_dp(qt), _dp(out))
    return out


# ---------------------------------------------------------------------------------
class RefRoutines:
    """The reference's own routines (oracle/_ref), driven on one cloud."""

    FIELD_IDS = dict(prim=1, q=2, dq=3, qm=4, delta=5, nx=6, ny=7, flux_res=8, temp=9)

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_LIB_PATH)

    def __init__(self, cloud, power=0.0, vl_const=150.0, interior_points_normal_flag=0,
                 format=1):
        self.lib = lib = C.CDLL(REF_LIB_PATH)
        lib.ref_setup.argtypes = [C.c_int32, C.c_int32] + [C.c_void_p] * 9 + \
            [C.c_double, C.c_double, C.c_int32, C.c_int32]
        lib.ref_set_split.argtypes = [C.c_int32] + [C.c_void_p] * 8
        lib.ref_field.argtypes = [C.c_int32, C.c_int32, C.c_void_p]
        lib.ref_qtilde_to_primitive.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_flux.argtypes = [C.c_int32] + [C.c_double] * 6 + [C.c_void_p]
        lib.ref_venkat.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
        self.n, self.w = cloud.n, cloud.max_nbhs
        self._keep = [np.ascontiguousarray(a) for a in (
            cloud.x.astype(np.float64), cloud.y.astype(np.float64),
            cloud.left.astype(np.int32), cloud.right.astype(np.int32),
            cloud.flag_1.astype(np.int32), cloud.flag_2.astype(np.int32),
            cloud.min_dist.astype(np.float64), cloud.nbhs.astype(np.int32),
            np.where(cloud.conn >= 0, cloud.conn, 0).astype(np.int32))]
        lib.ref_setup(self.n, self.w, *[a.ctypes.data for a in self._keep], power, vl_const,
                      interior_points_normal_flag, format)

    def set_split(self, orc: Oracle):
        arrs = [np.ascontiguousarray(orc.ifield(f).astype(np.int32)) for f in (
            "xpos_nbhs", "xneg_nbhs", "ypos_nbhs", "yneg_nbhs",
            "xpos_conn", "xneg_conn", "ypos_conn", "yneg_conn")]
        arrs = [np.where(a >= 0, a, 0).astype(np.int32) for a in arrs]
        self._split = arrs
        self.lib.ref_set_split(self.w, *[a.ctypes.data for a in arrs])

    def set(self, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self.lib.ref_field(self.FIELD_IDS[name], 0, a.ctypes.data)

    def get(self, name: str) -> np.ndarray:
        size = {"delta": 1, "nx": 1, "ny": 1, "dq": 8, "qm": 8, "temp": 8}.get(name, 4) * self.n
        a = np.zeros(size)
        self.lib.ref_field(self.FIELD_IDS[name], 1, a.ctypes.data)
        if size == 4 * self.n:
            return a.reshape(self.n, 4)
        if size == 8 * self.n:
            return a.reshape(self.n, 4, 2)
        return a

    def call(self, routine: str):
        getattr(self.lib, "ref_" + routine)()

    def qtilde_to_primitive(self, qt) -> np.ndarray:
        qt = np.ascontiguousarray(qt, dtype=np.float64)
        out = np.zeros(4)
        self.lib.ref_qtilde_to_primitive(qt.ctypes.data, out.ctypes.data)
        return out

    def flux(self, kind: int, nx, ny, u1, u2, rho, pr) -> np.ndarray:
        G = np.zeros(4)
        self.lib.ref_flux(kind, nx, ny, u1, u2, rho, pr, G.ctypes.data)
        return G

    def venkat(self, k0: int, qt) -> np.ndarray:
        """venkat_limiter for 0-based point k0 on the module's q, qm, min_dist."""
        qt = np.ascontiguousarray(qt, dtype=np.float64)
        phi = np.zeros(4)
        self.lib.ref_venkat(k0 + 1, qt.ctypes.data, phi.ctypes.data)
        return phi


# ---------------------------------------------------------------------------------
class DualRoutines:
    """The oracle in dual numbers, driven one routine at a time: set values and
    tangents of fields, call a routine, read the tangents of its outputs (the forward
    side of a dot-product check of a single reverse routine)."""

    def __init__(self, cloud, **params):
        lib = _load()
        prm = OrcParams(**default_params(**params))
        self.n, self.w = cloud.n, cloud.max_nbhs
        self._keep = [np.ascontiguousarray(a) for a in (
            cloud.x.astype(np.float64), cloud.y.astype(np.float64),
            cloud.left.astype(np.int32), cloud.right.astype(np.int32),
            cloud.flag_1.astype(np.int32), cloud.flag_2.astype(np.int32),
            cloud.min_dist.astype(np.float64), cloud.nbhs.astype(np.int32),
            np.where(cloud.conn >= 0, cloud.conn, 0).astype(np.int32),
            cloud.nx.astype(np.float64), cloud.ny.astype(np.float64))]
        x, y, l, r, f1, f2, md, nb, cn, nx, ny = self._keep
        self.h = lib.orc_dual_create(self.n, self.w, _dp(x), _dp(y), _ip(l), _ip(r), _ip(f1), _ip(f2),
                                     _dp(md), _ip(nb), _ip(cn), _dp(nx), _dp(ny), C.byref(prm))
        lib.orc_dual_setup(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            _load().orc_dual_destroy(self.h)
            self.h = None

    def set(self, name: str, value, tangent=None):
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(-1)
        t = None if tangent is None else np.ascontiguousarray(tangent, dtype=np.float64).reshape(-1)
        got = _load().orc_dual_set(self.h, name.encode(), _dp(v), None if t is None else _dp(t))
        if got != v.size:
            raise ValueError(f"{name}: size {v.size}, field holds {got}")

    def get(self, name: str):
        size = {"dq": 8, "qm": 8, "temp": 8, "prim": 4, "q": 4, "flux_res": 4}.get(name, 1) * self.n
        v, t = np.zeros(size), np.zeros(size)
        _load().orc_dual_get(self.h, name.encode(), _dp(v), _dp(t))
        return v, t

    def call(self, routine: str):
        if _load().orc_dual_call(self.h, routine.encode()) != 0:
            raise KeyError(routine)


TAP_LIB_PATH = os.path.join(HERE, "_tap", "libmfcfd_tap.so")


class TapRoutines:
    """The reference's Tapenade-generated reverse routines (oracle/_tap), driven on one
    cloud: point fields in, adjoint seeds in pointb, a reverse routine called, input
    adjoints read back.  Built only where the reference sources are present."""

    FIELD_IDS = dict(prim=1, q=2, dq=3, qm=4, delta=5, nx=6, ny=7, flux_res=8, x=10, y=11)
    SIZES = dict(prim=4, q=4, dq=8, qm=8, flux_res=4)

    @staticmethod
    def available() -> bool:
        return os.path.exists(TAP_LIB_PATH)

    def __init__(self, cloud, power=0.0, vl_const=150.0, interior_points_normal_flag=0, format=1):
        self.lib = lib = C.CDLL(TAP_LIB_PATH)
        lib.tap_setup.argtypes = [C.c_int32, C.c_int32] + [C.c_void_p] * 9 + \
            [C.c_double, C.c_double, C.c_int32, C.c_int32]
        lib.tap_set_split.argtypes = [C.c_int32] + [C.c_void_p] * 8
        lib.tap_field.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        lib.tap_flux_b.argtypes = [C.c_int32] + [C.c_double] * 6 + [C.c_void_p, C.c_void_p]
        lib.tap_qtilde_to_primitive_b.argtypes = [C.c_void_p] * 3
        lib.tap_qtilde_to_primitive.argtypes = [C.c_void_p] * 2
        lib.tap_venkat_b.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        self.n, self.w = cloud.n, cloud.max_nbhs
        self._keep = [np.ascontiguousarray(a) for a in (
            cloud.x.astype(np.float64), cloud.y.astype(np.float64),
            cloud.left.astype(np.int32), cloud.right.astype(np.int32),
            cloud.flag_1.astype(np.int32), cloud.flag_2.astype(np.int32),
            cloud.min_dist.astype(np.float64), cloud.nbhs.astype(np.int32),
            np.where(cloud.conn >= 0, cloud.conn, 0).astype(np.int32))]
        lib.tap_setup(self.n, self.w, *[a.ctypes.data for a in self._keep], power, vl_const,
                      interior_points_normal_flag, format)

    def set_split(self, orc: Oracle):
        arrs = [np.ascontiguousarray(orc.ifield(f).astype(np.int32)) for f in (
            "xpos_nbhs", "xneg_nbhs", "ypos_nbhs", "yneg_nbhs",
            "xpos_conn", "xneg_conn", "ypos_conn", "yneg_conn")]
        arrs = [np.where(a >= 0, a, 0).astype(np.int32) for a in arrs]
        self._split = arrs
        self.lib.tap_set_split(self.w, *[a.ctypes.data for a in arrs])

    def _io(self, name, adj, arr=None):
        size = self.SIZES.get(name, 1) * self.n
        if arr is None:
            a = np.zeros(size)
            self.lib.tap_field(self.FIELD_IDS[name], adj, 1, a.ctypes.data)
            return a
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        assert a.size == size, (name, a.size, size)
        self.lib.tap_field(self.FIELD_IDS[name], adj, 0, a.ctypes.data)

    def set(self, name, arr):
        self._io(name, 0, arr)

    def get(self, name):
        return self._io(name, 0)

    def set_adj(self, name, arr):
        self._io(name, 1, arr)

    def get_adj(self, name):
        return self._io(name, 1)

    def zero_adjoints(self):
        self.lib.tap_zero_adjoints()

    def call(self, routine: str):
        getattr(self.lib, "tap_" + routine)()

    def flux_b(self, kind, nx, ny, u1, u2, rho, pr, gb):
        """Adjoints (nx, ny, u1, u2, rho, pr) of one flux kind for output adjoint gb."""
        gb = np.ascontiguousarray(gb, dtype=np.float64)
        out = np.zeros(6)
        self.lib.tap_flux_b(kind, nx, ny, u1, u2, rho, pr, gb.ctypes.data, out.ctypes.data)
        return out

    def qtilde_to_primitive(self, qt):
        """(rho, u1, u2, pr) of q-tilde, by the differentiated module's forward."""
        qt = np.ascontiguousarray(qt, dtype=np.float64)
        out = np.zeros(4)
        self.lib.tap_qtilde_to_primitive(qt.ctypes.data, out.ctypes.data)
        return out

    def qtilde_to_primitive_b(self, qt, pb):
        """q-tilde adjoint for adjoints pb = (rho, u1, u2, pr) of the primitives."""
        qt = np.ascontiguousarray(qt, dtype=np.float64)
        pb = np.ascontiguousarray(pb, dtype=np.float64)
        out = np.zeros(4)
        self.lib.tap_qtilde_to_primitive_b(qt.ctypes.data, pb.ctypes.data, out.ctypes.data)
        return out

    def venkat_b(self, k0: int, qt, phib):
        """q-tilde adjoint of venkat_limiter at 0-based point k0; the q and qm adjoints
        of the point are added into pointb."""
        qt = np.ascontiguousarray(qt, dtype=np.float64)
        phib = np.ascontiguousarray(phib, dtype=np.float64)
        out = np.zeros(4)
        self.lib.tap_venkat_b(k0 + 1, qt.ctypes.data, phib.ctypes.data, out.ctypes.data)
        return out
