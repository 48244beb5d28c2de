"""mf_compute_cl_cd_cm_noalloc (no counterpart in the reference; for callers that
cannot allocate device memory): Cl, Cd, Cm and cp from caller-owned device arrays,
with the partial sums in the library's fixed device array, against
mf_compute_cl_cd_cm of a context holding the same state.  Two calls in a row, the
second on a smaller cloud (with two shapes), so the second reuses what the first
left in the fixed array."""
import dataclasses

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself links (inputs and outputs are the caller's
    device memory; the test allocates them here)."""
    import ctypes as C
    from mfcfd_amd import _lib
    _lib.lib()
    h = C.CDLL("libamdhip64.so.7")  # the soname libmfcfd_hip.so links: same handle
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


class _Dev:
    """One device array filled from (and read back into) a numpy array."""

    def __init__(self, hip, a):
        import ctypes as C
        self.hip, self.a = hip, np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.a.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(cloud, prim, nx, ny, shapes, mach, aoa):
    from mfcfd_amd import _lib
    hip = _hip()
    sidx = np.where(cloud.flag_2 > 0)[0]
    ins = [_Dev(hip, a) for a in (
        cloud.x.astype(np.float64), cloud.y.astype(np.float64), nx.astype(np.float64),
        ny.astype(np.float64), prim.reshape(-1).astype(np.float64), cloud.left.astype(np.int32),
        cloud.right.astype(np.int32), cloud.flag_2.astype(np.int32), sidx.astype(np.int32))]
    out = _Dev(hip, np.full(3 * shapes, np.nan))
    cp = _Dev(hip, np.full(len(sidx), np.nan))
    print(f"[no_alloc] {cloud.n} points, {len(sidx)} shape points, {shapes} shapes", flush=True)
    try:
        rc = _lib.lib().mf_compute_cl_cd_cm_noalloc(cloud.n, len(sidx), shapes, mach, aoa,
                                                    *[a.ptr for a in ins], out.ptr, cp.ptr, None)
        _lib.check(rc, "mf_compute_cl_cd_cm_noalloc")
        assert hip.hipDeviceSynchronize() == 0
        o = out.get().reshape(shapes, 3)
        cpv = cp.get()
    finally:
        for a in ins + [out, cp]:
            a.free()
    return o[:, 0], o[:, 1], o[:, 2], cpv


def _compare(cloud, params, iters, shapes=1):
    from mfcfd_amd.solver import LskumSolver
    with LskumSolver(cloud, shapes=shapes, **params) as g:
        g.q_lskum(iters)
        cl, cd, cm, cp = g.compute_cl_cd_cm()
        prim, nx, ny = g.prim, g.field("nx"), g.field("ny")
    ncl, ncd, ncm, ncp = _noalloc(cloud, prim, nx, ny, shapes, params["mach"], params["aoa"])
    for a, b in ((ncl, cl), (ncd, cd), (ncm, cm)):
        assert a.shape == b.shape
        assert rel_err(a, b) <= 1e-12, (a, b)
    assert rel_err(ncp, cp) <= 1e-15
    return cd


def test_noalloc_forces_two_calls(c1_cloud, small_cloud):
    p = dict(mach=0.63, aoa=2.0, cfl=0.2, vl_const=5.0, inner_iterations=0)
    cd1 = _compare(c1_cloud, p, 30)
    # second call: a smaller cloud, its lower surface relabelled as shape 2
    f2 = small_cloud.flag_2.copy()
    f2[(small_cloud.flag_1 == 0) & (small_cloud.y < 0.0)] = 2
    two = dataclasses.replace(small_cloud, flag_2=f2)
    cd2 = _compare(two, dict(p, mach=0.75, aoa=-1.0), 12, shapes=2)
    assert cd2.shape == (2,) and abs(cd1[0]) > 0


def test_noalloc_rejects_too_many_shapes(small_cloud):
    from mfcfd_amd import _lib
    rc = _lib.lib().mf_compute_cl_cd_cm_noalloc(small_cloud.n, 10, _lib.MF_NOALLOC_MAX_SHAPES + 1,
                                                0.6, 0.0, *([None] * 12))
    assert rc == _lib.MF_ERR_INVALID
    assert b"shapes" in _lib.lib().mf_last_error()
