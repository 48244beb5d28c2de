"""Point-cloud files in the reference's layouts.

* ``point/point.h5`` (point_preprocessor.F90:215-546): group ``/1`` with integer
  attributes ``total``, ``local``, ``ghost``; dataset ``/1/local`` of float64, C order
  ``[local][12 + max nbhs]`` with columns original_id, x, y, nx, ny, min_dist, left,
  right, qtdepth, flag_1, flag_2, nbhs, conn(1..nbhs) (1-based; integers stored as
  doubles, :357-371); dataset ``/1/ghost`` ``[ghost][4]`` (opened unconditionally,
  :439-442; empty for one partition).
* legacy text ``point/partGrid`` (:23-146): header ``nproc max_points local_points
  ghost_points``, then per point ``id x y left right flag_1 flag_2 min_dist nbhs
  conn...`` (format 1) or ``id x y left right flag_1 flag_2 nx ny qtdepth min_dist
  nbhs conn...`` (format 2, "quadtree").

h5py is not part of this environment, so the HDF5 files are written and read through
the HDF5 C library (libhdf5 1.10) with ctypes.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

from .cloud import PointCloud

_H5 = None
_hid = C.c_int64
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5F_ACC_RDONLY = 0
_H5F_ACC_TRUNC = 2
_H5S_SCALAR = 0


def _candidates():
    env = os.environ.get("MFCFD_HDF5_LIB")
    if env:
        yield env
    yield "/opt/conda/lib/libhdf5.so"
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found


def h5lib():
    """The HDF5 C library (raises if none can be loaded)."""
    global _H5
    if _H5 is not None:
        return _H5
    err = None
    for path in _candidates():
        try:
            lib = C.CDLL(path)
            break
        except OSError as e:  # try the next candidate
            err = e
    else:
        raise RuntimeError(f"libhdf5 not found ({err}); set MFCFD_HDF5_LIB")
    sig = {
        "H5open": (C.c_int, []),
        "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
        "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
        "H5Fclose": (C.c_int, [_hid]),
        "H5Gcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid]),
        "H5Gopen2": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Gclose": (C.c_int, [_hid]),
        "H5Screate": (_hid, [C.c_int]),
        "H5Screate_simple": (_hid, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
        "H5Sclose": (C.c_int, [_hid]),
        "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Dwrite": (C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dread": (C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dclose": (C.c_int, [_hid]),
        "H5Acreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]),
        "H5Aopen": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Awrite": (C.c_int, [_hid, _hid, C.c_void_p]),
        "H5Aread": (C.c_int, [_hid, _hid, C.c_void_p]),
        "H5Aclose": (C.c_int, [_hid]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.H5open() < 0:
        raise RuntimeError("H5open failed")
    _H5 = lib
    return lib


def _type(name: str) -> int:
    return _hid.in_dll(h5lib(), name).value


def _ok(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return v


def _dims(*d):
    return (C.c_uint64 * len(d))(*d)


def local_table(cloud: PointCloud, original_id: np.ndarray | None = None) -> np.ndarray:
    """The ``[N][12 + W]`` float64 table of ``/1/local`` (1-based ids and links)."""
    n, w = cloud.n, cloud.max_nbhs
    t = np.zeros((n, 12 + w), dtype=np.float64)
    t[:, 0] = np.arange(1, n + 1) if original_id is None else original_id
    t[:, 1], t[:, 2] = cloud.x, cloud.y
    t[:, 3], t[:, 4] = cloud.nx, cloud.ny
    t[:, 5] = cloud.min_dist
    t[:, 6] = np.where(cloud.left >= 0, cloud.left + 1, 0)
    t[:, 7] = np.where(cloud.right >= 0, cloud.right + 1, 0)
    t[:, 8] = cloud.qtdepth
    t[:, 9], t[:, 10], t[:, 11] = cloud.flag_1, cloud.flag_2, cloud.nbhs
    t[:, 12:] = np.where(cloud.conn >= 0, cloud.conn + 1, 0)
    return t


def cloud_from_table(t: np.ndarray) -> PointCloud:
    """Inverse of local_table (0-based arrays, -1 padding)."""
    n = t.shape[0]
    w = t.shape[1] - 12
    nbhs = t[:, 11].astype(np.int32)
    conn = t[:, 12:].astype(np.int64) - 1
    conn[np.arange(w)[None, :] >= nbhs[:, None]] = -1
    return PointCloud(x=t[:, 1].copy(), y=t[:, 2].copy(), nx=t[:, 3].copy(), ny=t[:, 4].copy(),
                      min_dist=t[:, 5].copy(), left=t[:, 6].astype(np.int32) - 1,
                      right=t[:, 7].astype(np.int32) - 1, flag_1=t[:, 9].astype(np.int32),
                      flag_2=t[:, 10].astype(np.int32), nbhs=nbhs, conn=conn.astype(np.int32),
                      qtdepth=t[:, 8].astype(np.int32), meta={"original_id": t[:, 0].astype(np.int64)})


def write_point_h5(path: str, cloud: PointCloud) -> None:
    """Write ``cloud`` as a one-partition point.h5 (group /1)."""
    h5 = h5lib()
    t = np.ascontiguousarray(local_table(cloud))
    f = _ok(h5.H5Fcreate(path.encode(), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), "H5Fcreate")
    try:
        g = _ok(h5.H5Gcreate2(f, b"/1", _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT), "H5Gcreate2")
        i32, f64 = _type("H5T_STD_I32LE_g"), _type("H5T_IEEE_F64LE_g")
        nint, ndbl = _type("H5T_NATIVE_INT_g"), _type("H5T_NATIVE_DOUBLE_g")
        for name, val in ((b"total", cloud.n), (b"local", cloud.n), (b"ghost", 0)):
            sp = _ok(h5.H5Screate(_H5S_SCALAR), "H5Screate")
            a = _ok(h5.H5Acreate2(g, name, i32, sp, _H5P_DEFAULT, _H5P_DEFAULT), "H5Acreate2")
            v = C.c_int32(int(val))
            _ok(h5.H5Awrite(a, nint, C.byref(v)), "H5Awrite")
            h5.H5Aclose(a)
            h5.H5Sclose(sp)
        for name, arr in ((b"local", t), (b"ghost", np.zeros((0, 4)))):
            sp = _ok(h5.H5Screate_simple(2, _dims(*arr.shape), None), "H5Screate_simple")
            d = _ok(h5.H5Dcreate2(g, name, f64, sp, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT), "H5Dcreate2")
            if arr.size:
                _ok(h5.H5Dwrite(d, ndbl, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, arr.ctypes.data), "H5Dwrite")
            h5.H5Dclose(d)
            h5.H5Sclose(sp)
        h5.H5Gclose(g)
    finally:
        h5.H5Fclose(f)


def read_point_h5(path: str) -> PointCloud:
    """Read group /1 of a point.h5 (as read_hdf5input_point_data does)."""
    h5 = h5lib()
    f = _ok(h5.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen")
    try:
        g = _ok(h5.H5Gopen2(f, b"/1", _H5P_DEFAULT), "H5Gopen2")
        attrs = {}
        for name in (b"total", b"local", b"ghost"):
            a = _ok(h5.H5Aopen(g, name, _H5P_DEFAULT), "H5Aopen")
            v = C.c_int32()
            _ok(h5.H5Aread(a, _type("H5T_NATIVE_INT_g"), C.byref(v)), "H5Aread")
            attrs[name.decode()] = v.value
            h5.H5Aclose(a)
        d = _ok(h5.H5Dopen2(g, b"local", _H5P_DEFAULT), "H5Dopen2")
        sp = h5.H5Dget_space(d)
        dims = (C.c_uint64 * 2)()
        _ok(h5.H5Sget_simple_extent_dims(sp, dims, None), "H5Sget_simple_extent_dims")
        t = np.empty((dims[0], dims[1]), dtype=np.float64)
        _ok(h5.H5Dread(d, _type("H5T_NATIVE_DOUBLE_g"), _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, t.ctypes.data), "H5Dread")
        h5.H5Sclose(sp)
        h5.H5Dclose(d)
        h5.H5Gclose(g)
    finally:
        h5.H5Fclose(f)
    if attrs["local"] != t.shape[0]:
        raise RuntimeError(f"{path}: local = {attrs['local']} but /1/local has {t.shape[0]} rows")
    return cloud_from_table(t[: attrs["local"]])


def write_partgrid(path: str, cloud: PointCloud, fmt: int = 1) -> None:
    """Legacy text point file (format 1, or 2 with nx ny qtdepth)."""
    t = local_table(cloud)
    with open(path, "w") as fh:
        fh.write(f"1 {cloud.n} {cloud.n} 0\n")
        for row in t:
            nb = int(row[11])
            conn = " ".join(str(int(c)) for c in row[12:12 + nb])
            head = f"{int(row[0])} {row[1]:.17g} {row[2]:.17g} {int(row[6])} {int(row[7])} {int(row[9])} {int(row[10])}"
            if fmt == 2:
                head += f" {row[3]:.17g} {row[4]:.17g} {int(row[8])}"
            fh.write(f"{head} {row[5]:.17g} {nb} {conn}\n")


def read_partgrid(path: str, fmt: int = 1) -> PointCloud:
    with open(path) as fh:
        nproc, max_points, local_points, ghost = (int(v) for v in fh.readline().split())
        if nproc != 1:
            raise RuntimeError(f"{path}: {nproc} partitions (one expected)")
        rows = [fh.readline().split() for _ in range(local_points)]
    w = max(int(r[11 if fmt == 2 else 8]) for r in rows)
    t = np.zeros((local_points, 12 + w))
    for k, r in enumerate(rows):
        t[k, 0], t[k, 1], t[k, 2] = float(r[0]), float(r[1]), float(r[2])
        t[k, 6], t[k, 7], t[k, 9], t[k, 10] = float(r[3]), float(r[4]), float(r[5]), float(r[6])
        j = 7
        if fmt == 2:
            t[k, 3], t[k, 4], t[k, 8] = float(r[7]), float(r[8]), float(r[9])
            j = 10
        t[k, 5] = float(r[j])
        nb = int(r[j + 1])
        t[k, 11] = nb
        t[k, 12:12 + nb] = [float(v) for v in r[j + 2:j + 2 + nb]]
    return cloud_from_table(t)


def write_case(path: str, **opts) -> None:
    """A case.in of PETSc-style ``-key value`` lines (examples/mpi_primal/case.in.example)."""
    with open(path, "w") as fh:
        for k, v in opts.items():
            fh.write(f"-{k} {v}\n")
