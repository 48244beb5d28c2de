"""Point files and the Fortran host's input side, on the CPU.

point.h5 (point_preprocessor.F90:215-546) and the legacy partGrid text (:23-146) are
written by mfcfd_amd/pointio.py and read back both by Python and by the Fortran host
program (``meshfree_host --check`` reads case.in and the point file and exits before
any device call)."""
import os
import re
import subprocess

import numpy as np
import pytest

from mfcfd_amd import pointio
from mfcfd_amd.build import HOST

FIELDS = ("x", "y", "min_dist", "left", "right", "flag_1", "flag_2", "nbhs", "conn")


def _same(a, b, extra=()):
    for f in FIELDS + tuple(extra):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_point_h5_round_trip(small_cloud, tmp_path):
    p = str(tmp_path / "point.h5")
    pointio.write_point_h5(p, small_cloud)
    _same(small_cloud, pointio.read_point_h5(p), ("nx", "ny"))


@pytest.mark.parametrize("fmt", [1, 2])
def test_partgrid_round_trip(small_cloud, tmp_path, fmt):
    p = str(tmp_path / "partGrid")
    pointio.write_partgrid(p, small_cloud, fmt)
    _same(small_cloud, pointio.read_partgrid(p, fmt), ("nx", "ny") if fmt == 2 else ())


def _host_check(case_dir):
    if not os.path.exists(HOST):
        pytest.skip("meshfree_host not built (amdflang or libhdf5 missing)")
    out = subprocess.run([HOST, "--check"], cwd=case_dir, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("check:")][0]
    kv = dict(re.findall(r"(\w+)=\s*(\S+)", line))
    return int(kv["n"]), int(kv["w"]), int(kv["sum_nbhs"]), float(kv["sum_xy"])


@pytest.mark.parametrize("kind", ["h5", "partgrid1", "partgrid2"])
def test_fortran_host_reads_points(small_cloud, tmp_path, kind):
    c = small_cloud
    (tmp_path / "point").mkdir()
    if kind == "h5":
        pointio.write_point_h5(str(tmp_path / "point" / "point.h5"), c)
        fmt = "legacy"
    else:
        f = int(kind[-1])
        pointio.write_partgrid(str(tmp_path / "point" / "partGrid"), c, f)
        fmt = "legacy" if f == 1 else "quadtree"
    pointio.write_case(str(tmp_path / "case.in"), cfl=0.2, max_iters=5, format_file=fmt)
    n, w, snb, sxy = _host_check(str(tmp_path))
    assert (n, w, snb) == (c.n, int(c.nbhs.max()), int(c.nbhs.sum()))
    assert abs(sxy - (c.x.sum() + c.y.sum())) <= 1e-12 * abs(sxy)


def test_fortran_host_rejects_bad_shapes(small_cloud, tmp_path):
    """flag_2 above -shapes aborts, as point_preprocessor.F90:383-385 does."""
    if not os.path.exists(HOST):
        pytest.skip("meshfree_host not built")
    (tmp_path / "point").mkdir()
    c = small_cloud
    bad = pointio.cloud_from_table(pointio.local_table(c))
    bad.flag_2[:3] = 2
    pointio.write_point_h5(str(tmp_path / "point" / "point.h5"), bad)
    pointio.write_case(str(tmp_path / "case.in"), cfl=0.2, shapes=1)
    out = subprocess.run([HOST, "--check"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert out.returncode != 0 and "shapes value wrong" in out.stdout
