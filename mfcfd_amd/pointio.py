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
* partitioned files for an N-rank run (:18-22, :218): group ``/r`` (r = rank + 1) of
  point.h5, or text file ``point/partGrid`` + the rank in 4 digits.  A rank's local
  points are numbered 1..local and its ghosts local+1..total; conn, left and right use
  that numbering; the ghost rows are (PETSc global index, 1-based; x; y; min_dist)
  (:450-461, :128-135), the global index of a point being its rank's offset (the local
  counts of the lower ranks) plus its local index (petsc_data_structure.F90:26-31).

h5py is not part of this environment, so the HDF5 files are written and read through
the HDF5 C library (libhdf5 1.10) with ctypes.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import dataclasses
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


def _hilbert_index(x: np.ndarray, y: np.ndarray, order: int = 16) -> np.ndarray:
    """Hilbert-curve index of each point on a 2^order grid over the bounding box."""
    g = (1 << order) - 1
    span = lambda a: (a - a.min()) / max(float(a.max() - a.min()), 1e-300)
    xi = np.rint(span(x) * g).astype(np.int64)
    yi = np.rint(span(y) * g).astype(np.int64)
    d = np.zeros_like(xi)
    s = 1 << (order - 1)
    while s > 0:
        rx = (xi & s) > 0
        ry = (yi & s) > 0
        d += s * s * ((3 * rx.astype(np.int64)) ^ ry.astype(np.int64))
        flip = ~ry & rx
        xi[flip] = g - xi[flip]
        yi[flip] = g - yi[flip]
        swap = ~ry
        xi[swap], yi[swap] = yi[swap], xi[swap].copy()
        s >>= 1
    return d


@dataclasses.dataclass
class Partition:
    """One rank's part of a partitioned point file."""

    local: np.ndarray  # [local][12 + W] table in the rank's numbering (see local_table)
    ghost: np.ndarray  # [ghost][4]: PETSc global index (1-based), x, y, min_dist

    @property
    def total(self) -> int:
        return self.local.shape[0] + self.ghost.shape[0]


def partition(cloud: PointCloud, nparts: int) -> list[Partition]:
    """Split ``cloud`` into ``nparts`` ranks of contiguous Hilbert-curve ranges (equal
    point counts), each with its 1-ring ghosts and the points its boundary points name
    as left/right.  Column 1 of each local table is the point's 1-based id in ``cloud``."""
    n = cloud.n
    if not 1 <= nparts <= n:
        raise ValueError(f"nparts must be in 1..{n}")
    order = np.argsort(_hilbert_index(cloud.x, cloud.y), kind="stable")
    bounds = np.linspace(0, n, nparts + 1).round().astype(np.int64)
    rank_of = np.empty(n, np.int64)
    pidx = np.empty(n, np.int64)  # PETSc global index (0-based): rank blocks in order
    for r in range(nparts):
        pts = order[bounds[r]:bounds[r + 1]]
        rank_of[pts] = r
        pidx[pts] = bounds[r] + np.arange(len(pts))
    full = local_table(cloud)
    w = cloud.max_nbhs
    parts = []
    for r in range(nparts):
        pts = order[bounds[r]:bounds[r + 1]]
        refs = [cloud.conn[pts][cloud.conn[pts] >= 0], cloud.left[pts][cloud.left[pts] >= 0],
                cloud.right[pts][cloud.right[pts] >= 0]]
        ext = np.unique(np.concatenate(refs))
        ghosts = ext[rank_of[ext] != r]
        ghosts = ghosts[np.argsort(pidx[ghosts], kind="stable")]
        num = np.full(n, -1, np.int64)  # global point -> 1-based number on this rank
        num[pts] = np.arange(1, len(pts) + 1)
        num[ghosts] = len(pts) + np.arange(1, len(ghosts) + 1)
        t = full[pts].copy()
        for col in (6, 7):
            ref = t[:, col].astype(np.int64)
            t[:, col] = np.where(ref > 0, num[np.maximum(ref - 1, 0)], 0)
        c = t[:, 12:12 + w].astype(np.int64)
        t[:, 12:12 + w] = np.where(c > 0, num[np.maximum(c - 1, 0)], 0)
        g = np.stack([pidx[ghosts] + 1.0, cloud.x[ghosts], cloud.y[ghosts], cloud.min_dist[ghosts]], axis=1) \
            if len(ghosts) else np.zeros((0, 4))
        parts.append(Partition(local=t, ghost=g))
    return parts


def assemble(parts: list[Partition]) -> PointCloud:
    """The global cloud of a partitioned file, in original-id order (inverse of
    ``partition``; ghost references resolved through the PETSc global numbering)."""
    counts = [p.local.shape[0] for p in parts]
    offs = np.concatenate([[0], np.cumsum(counts)])
    ids = np.concatenate([p.local[:, 0] for p in parts]).astype(np.int64)  # by PETSc index
    n = len(ids)
    if not np.array_equal(np.sort(ids), np.arange(1, n + 1)):
        raise RuntimeError("partitioned file: local original ids are not a permutation of 1..N")
    w = max(p.local.shape[1] for p in parts) - 12
    out = np.zeros((n, 12 + w))
    for r, p in enumerate(parts):
        L = p.local.shape[0]
        # this rank's numbering -> original id (1-based)
        to_id = np.concatenate([[0], ids[offs[r]:offs[r] + L], ids[p.ghost[:, 0].astype(np.int64) - 1]])
        t = p.local
        rows = t[:, 0].astype(np.int64) - 1
        out[rows, :12] = t[:, :12]
        for col in (6, 7):
            out[rows, col] = to_id[t[:, col].astype(np.int64)]
        wr = t.shape[1] - 12
        out[rows, 12:12 + wr] = to_id[t[:, 12:].astype(np.int64)]
    return cloud_from_table(out)


def write_point_h5(path: str, cloud: PointCloud, nparts: int = 1) -> None:
    """Write ``cloud`` as point.h5: one partition (group /1, original order) or
    ``nparts`` ranks (groups /1../nparts, see ``partition``)."""
    parts = [Partition(local_table(cloud), np.zeros((0, 4)))] if nparts == 1 else partition(cloud, nparts)
    h5 = h5lib()
    f = _ok(h5.H5Fcreate(path.encode(), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), "H5Fcreate")
    try:
        for r, part in enumerate(parts):
            _write_group(h5, f, r + 1, part)
    finally:
        h5.H5Fclose(f)


def _write_group(h5, f, gid: int, part: Partition) -> None:
    g = _ok(h5.H5Gcreate2(f, f"/{gid}".encode(), _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT), "H5Gcreate2")
    i32, f64 = _type("H5T_STD_I32LE_g"), _type("H5T_IEEE_F64LE_g")
    nint, ndbl = _type("H5T_NATIVE_INT_g"), _type("H5T_NATIVE_DOUBLE_g")
    for name, val in ((b"total", part.total), (b"local", part.local.shape[0]), (b"ghost", part.ghost.shape[0])):
        sp = _ok(h5.H5Screate(_H5S_SCALAR), "H5Screate")
        a = _ok(h5.H5Acreate2(g, name, i32, sp, _H5P_DEFAULT, _H5P_DEFAULT), "H5Acreate2")
        v = C.c_int32(int(val))
        _ok(h5.H5Awrite(a, nint, C.byref(v)), "H5Awrite")
        h5.H5Aclose(a)
        h5.H5Sclose(sp)
    for name, arr in ((b"local", part.local), (b"ghost", part.ghost)):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        sp = _ok(h5.H5Screate_simple(2, _dims(*arr.shape), None), "H5Screate_simple")
        d = _ok(h5.H5Dcreate2(g, name, f64, sp, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT), "H5Dcreate2")
        if arr.size:
            _ok(h5.H5Dwrite(d, ndbl, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, arr.ctypes.data), "H5Dwrite")
        h5.H5Dclose(d)
        h5.H5Sclose(sp)
    h5.H5Gclose(g)


def _read_dataset(h5, g, name: bytes) -> np.ndarray:
    d = _ok(h5.H5Dopen2(g, name, _H5P_DEFAULT), "H5Dopen2")
    sp = h5.H5Dget_space(d)
    dims = (C.c_uint64 * 2)()
    _ok(h5.H5Sget_simple_extent_dims(sp, dims, None), "H5Sget_simple_extent_dims")
    t = np.empty((dims[0], dims[1]), dtype=np.float64)
    if t.size:
        _ok(h5.H5Dread(d, _type("H5T_NATIVE_DOUBLE_g"), _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, t.ctypes.data),
            "H5Dread")
    h5.H5Sclose(sp)
    h5.H5Dclose(d)
    return t


def _read_group(h5, f, gid: int, path: str) -> Partition:
    g = _ok(h5.H5Gopen2(f, f"/{gid}".encode(), _H5P_DEFAULT), "H5Gopen2")
    try:
        attrs = {}
        for name in (b"total", b"local", b"ghost"):
            a = _ok(h5.H5Aopen(g, name, _H5P_DEFAULT), "H5Aopen")
            v = C.c_int32()
            _ok(h5.H5Aread(a, _type("H5T_NATIVE_INT_g"), C.byref(v)), "H5Aread")
            attrs[name.decode()] = v.value
            h5.H5Aclose(a)
        t = _read_dataset(h5, g, b"local")
        gh = _read_dataset(h5, g, b"ghost")
    finally:
        h5.H5Gclose(g)
    if attrs["local"] != t.shape[0]:
        raise RuntimeError(f"{path}: local = {attrs['local']} but /{gid}/local has {t.shape[0]} rows")
    if attrs["ghost"] != gh.shape[0] or attrs["total"] != attrs["local"] + attrs["ghost"]:
        raise RuntimeError(f"{path}: /{gid} attributes {attrs} disagree with {gh.shape[0]} ghost rows")
    return Partition(local=t, ghost=gh)


def read_point_h5(path: str) -> PointCloud:
    """Read group /1 of a point.h5 (as read_hdf5input_point_data does at one rank)."""
    return cloud_from_table(read_point_h5_parts(path, 1)[0].local)


def read_point_h5_parts(path: str, nparts: int) -> list[Partition]:
    """Groups /1../nparts of a partitioned point.h5 (``assemble`` joins them)."""
    h5 = h5lib()
    f = _ok(h5.H5Fopen(path.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen")
    try:
        return [_read_group(h5, f, r + 1, path) for r in range(nparts)]
    finally:
        h5.H5Fclose(f)


def partgrid_name(path: str, rank: int, nparts: int) -> str:
    """``point/partGrid`` at one rank, ``point/partGrid0003`` for rank 3 of several
    (point_preprocessor.F90:18-19, itos(4, rank))."""
    return path if nparts == 1 else f"{path}{rank:04d}"


def write_partgrid(path: str, cloud: PointCloud, fmt: int = 1, nparts: int = 1) -> None:
    """Legacy text point file(s) (format 1, or 2 with nx ny qtdepth): one file at one
    rank, else one file per rank with its ghost rows after the local ones."""
    parts = [Partition(local_table(cloud), np.zeros((0, 4)))] if nparts == 1 else partition(cloud, nparts)
    for r, part in enumerate(parts):
        with open(partgrid_name(path, r, nparts), "w") as fh:
            fh.write(f"{nparts} {part.total} {part.local.shape[0]} {part.ghost.shape[0]}\n")
            for row in part.local:
                nb = int(row[11])
                conn = " ".join(str(int(c)) for c in row[12:12 + nb])
                head = f"{int(row[0])} {row[1]:.17g} {row[2]:.17g} {int(row[6])} {int(row[7])} {int(row[9])} {int(row[10])}"
                if fmt == 2:
                    head += f" {row[3]:.17g} {row[4]:.17g} {int(row[8])}"
                fh.write(f"{head} {row[5]:.17g} {nb} {conn}\n")
            for row in part.ghost:
                fh.write(f"{int(row[0])} {row[1]:.17g} {row[2]:.17g} {row[3]:.17g}\n")


def _read_partgrid_file(path: str, fmt: int, nparts: int) -> Partition:
    with open(path) as fh:
        nproc, max_points, local_points, ghost = (int(v) for v in fh.readline().split())
        if nproc != nparts:
            raise RuntimeError(f"{path}: {nproc} partitions ({nparts} expected)")
        if max_points != local_points + ghost:
            raise RuntimeError(f"{path}: header {max_points} != {local_points} + {ghost}")
        rows = [fh.readline().split() for _ in range(local_points)]
        grows = [fh.readline().split() for _ in range(ghost)]
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
    g = np.array([[float(v) for v in r[:4]] for r in grows]) if ghost else np.zeros((0, 4))
    return Partition(local=t, ghost=g)


def read_partgrid(path: str, fmt: int = 1) -> PointCloud:
    """One-rank legacy text point file."""
    return cloud_from_table(_read_partgrid_file(path, fmt, 1).local)


def read_partgrid_parts(path: str, nparts: int, fmt: int = 1) -> list[Partition]:
    """The per-rank legacy text files of an nparts-rank run (``assemble`` joins them)."""
    return [_read_partgrid_file(partgrid_name(path, r, nparts), fmt, nparts) for r in range(nparts)]


def write_case(path: str, **opts) -> None:
    """A case.in of PETSc-style ``-key value`` lines (examples/mpi_primal/case.in.example)."""
    with open(path, "w") as fh:
        for k, v in opts.items():
            fh.write(f"-{k} {v}\n")
