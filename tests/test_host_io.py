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


@pytest.mark.parametrize("nparts", [2, 5])
def test_partitioned_point_files_round_trip(small_cloud, tmp_path, nparts):
    """point.h5 groups /1../N and partGridNNNN files of an N-rank run: each rank's local
    points, ghosts (PETSc global index, x, y, min_dist) and local-numbered conn/left/right
    reassemble into the original cloud."""
    from mfcfd_amd import pointio
    c = small_cloud
    ref = pointio.local_table(c)
    parts = pointio.partition(c, nparts)
    assert sum(p.local.shape[0] for p in parts) == c.n
    offs = np.concatenate([[0], np.cumsum([p.local.shape[0] for p in parts])])
    owner_ids = np.concatenate([p.local[:, 0] for p in parts]).astype(int)  # by PETSc index
    for r, p in enumerate(parts):
        L = p.local.shape[0]
        # every reference past the local block is a ghost, and every ghost is referenced
        refs = np.concatenate([p.local[:, 12:].ravel(), p.local[:, 6], p.local[:, 7]]).astype(int)
        refs = refs[refs > 0]
        assert refs.max() <= p.total
        assert set(refs[refs > L]) == set(range(L + 1, p.total + 1))
        # ghosts carry the owner's coordinates, and no ghost is owned by this rank
        gid = p.ghost[:, 0].astype(int)
        assert np.all((gid - 1 < offs[r]) | (gid - 1 >= offs[r + 1]))
        orig = owner_ids[gid - 1] - 1
        assert np.array_equal(p.ghost[:, 1], c.x[orig]) and np.array_equal(p.ghost[:, 2], c.y[orig])
        assert np.array_equal(p.ghost[:, 3], c.min_dist[orig])
    back = pointio.local_table(pointio.assemble(parts))
    assert np.array_equal(back[:, 1:], ref[:, 1:])
    pointio.write_point_h5(str(tmp_path / "point.h5"), c, nparts=nparts)
    h5 = pointio.assemble(pointio.read_point_h5_parts(str(tmp_path / "point.h5"), nparts))
    assert np.array_equal(pointio.local_table(h5)[:, 1:], ref[:, 1:])
    pointio.write_partgrid(str(tmp_path / "partGrid"), c, fmt=2, nparts=nparts)
    assert (tmp_path / f"partGrid{nparts - 1:04d}").exists()
    txt = pointio.assemble(pointio.read_partgrid_parts(str(tmp_path / "partGrid"), nparts, fmt=2))
    assert np.array_equal(pointio.local_table(txt)[:, 1:], ref[:, 1:])


INPUT_NML = """&input_parameters
	cfl = 0.2
	max_iters = 7
	mach = 0.63
	vl_const = 650
	format_file = 'legacy'
	tscheme = 'ssprk43'
	blockx = 32
	inner_iterations = 1
	aoa = 2
/
"""


def test_fortran_host_reads_input_nml(small_cloud, tmp_path):
    """Without a case.in the host takes the CUDA build's input.nml
    (examples/cuda/input.nml.example's keys) and echoes the namelist."""
    if not os.path.exists(HOST):
        pytest.skip("meshfree_host not built")
    (tmp_path / "point").mkdir()
    pointio.write_point_h5(str(tmp_path / "point" / "point.h5"), small_cloud)
    (tmp_path / "input.nml").write_text(INPUT_NML)
    out = subprocess.run([HOST, "--check"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    echo = out.stdout.upper().replace(" ", "")
    assert "MAX_ITERS=7" in echo and "BLOCKX=32" in echo and "VL_CONST=650" in echo
    assert re.search(r"tscheme\s*:\s*ssprk43", out.stdout)
    assert f"check: n={small_cloud.n}" in out.stdout
    (tmp_path / "input.nml").write_text(INPUT_NML.replace("tscheme = 'ssprk43'", "tscheme = 'bogus'"))
    out = subprocess.run([HOST, "--check"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert out.returncode != 0 and "tscheme" in out.stdout
