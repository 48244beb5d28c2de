"""The Fortran host program end to end on the GPU: case.in + point/point.h5 in,
residue, solution/sol.dat, cp/cp-file and sensitivity/sensitivity.dat out, checked
against the Python binding of the same library and against the FP64 oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import rel_err
from mfcfd_amd import pointio
from mfcfd_amd.build import HOST

pytestmark = pytest.mark.gpu

CASE = dict(cfl=0.2, mach=0.63, aoa=2.0, vl_const=5.0, inner_iterations=2)


def _case(tmp_path, cloud, **opts):
    (tmp_path / "point").mkdir(exist_ok=True)
    pointio.write_point_h5(str(tmp_path / "point" / "point.h5"), cloud)
    o = dict(CASE)
    o.update(opts)
    o.setdefault("autotune", "no")
    pointio.write_case(str(tmp_path / "case.in"), **o)


def _run(case_dir):
    if not os.path.exists(HOST):
        pytest.fail("meshfree_host was not built")
    print(f"[host] running in {case_dir}", flush=True)
    out = subprocess.run([HOST], cwd=str(case_dir), capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-3000:] + out.stderr[-3000:]
    return out.stdout


def _read_sol(path):
    with open(path) as fh:
        head = fh.readline().split()
        rows = np.loadtxt(fh)
    return (int(head[0]), int(head[1]), float(head[2])), rows


def test_host_primal_matches_binding_and_oracle(small_cloud, tmp_path):
    from mfcfd_amd.solver import LskumSolver
    from oracle.oracle import Oracle
    _case(tmp_path, small_cloud, max_iters=20, nsave=10)
    out = _run(tmp_path)
    assert sum(1 for ln in out.splitlines() if ln.strip().startswith("iterations:")) == 20
    res = np.loadtxt(tmp_path / "residue")
    (n, itm1, res_old), sol = _read_sol(tmp_path / "solution" / "sol.dat")
    cpf = np.loadtxt(tmp_path / "cp" / "cp-file")
    with LskumSolver(small_cloud, **CASE) as s:
        r, rn, _ = s.q_lskum(20)
        prim = s.prim
        _, cd, _, cp = s.compute_cl_cd_cm()
    assert np.array_equal(res[:, 0], np.arange(1, 21))
    assert np.abs(res[:, 1] - r).max() <= 1e-13
    assert (n, itm1) == (small_cloud.n, 19) and abs(res_old - rn[1]) <= 1e-15 * rn[1]
    assert np.array_equal(sol[:, 0], np.arange(1, small_cloud.n + 1))
    assert rel_err(sol[:, 5:9], prim) <= 1e-15
    assert rel_err(cpf[:, 2], cp) <= 1e-13
    o = Oracle(small_cloud, aoa_deg=CASE["aoa"], **{k: v for k, v in CASE.items() if k != "aoa"}).setup()
    o.run(20)
    assert rel_err(sol[:, 5:9], o.field("prim")) < 1e-10


def test_host_restart_continues(small_cloud, tmp_path):
    """10 iterations, then a restart from solution/sol.dat for 10 more, end where 20
    straight iterations end (restart/sol.dat: initial_conditions.F90:44-69)."""
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    _case(a, small_cloud, max_iters=10, nsave=10)
    _run(a)
    _case(b, small_cloud, max_iters=10, restart_solution="same")
    (b / "restart").mkdir()
    shutil.copy(a / "solution" / "sol.dat", b / "restart" / "sol.dat")
    _run(b)
    _, sol = _read_sol(b / "solution" / "sol.dat")
    res = np.loadtxt(b / "residue")
    # the saved header is "it - 1" (9), so the restarted run numbers from 10
    assert res[0, 0] == 10 and res[-1, 0] == 19
    from mfcfd_amd.solver import LskumSolver
    with LskumSolver(small_cloud, **CASE) as s:
        s.q_lskum(20)
        prim = s.prim
    assert rel_err(sol[:, 5:9], prim) <= 1e-14


@pytest.mark.parametrize("mode", ["checkpoints", "tape"])
def test_host_adjoint_writes_sensitivity(small_cloud, tmp_path, mode):
    """-adjoint_mode checkpoints -chkpts 4 (binomial), and -adjoint_mode tape (every
    iteration taped in the primal sweep): the same sensitivities as mf_adjoint."""
    from mfcfd_amd.solver import LskumSolver
    opts = dict(chkpts=4) if mode == "checkpoints" else {}
    _case(tmp_path, small_cloud, max_iters=12, adjoint_mode=mode, **opts)
    out = _run(tmp_path)
    assert "Objective" in out
    with open(tmp_path / "sensitivity" / "sensitivity.dat") as fh:
        n, itm1 = (int(v) for v in fh.readline().split())
        sens = np.loadtxt(fh)
    assert (n, itm1) == (small_cloud.n, 12)
    with LskumSolver(small_cloud, **CASE) as s:
        J, dx, dy, _ = s.adjoint(12, snaps=4)
    g = np.concatenate([dx, dy])
    assert np.abs(np.concatenate([sens[:, 1], sens[:, 2]]) - g).max() <= 1e-10 * np.abs(g).max()


def test_host_forward_mode(small_cloud, tmp_path):
    """-adjoint_mode forward: src_mpi_forward's tangent (q_lskum_d.F90:34-39 seeds x of
    the point with original id 1), the printed derivative equal to mf_tangent's and to
    the adjoint's dJ/dx of that point."""
    import re
    from mfcfd_amd.solver import LskumSolver
    _case(tmp_path, small_cloud, max_iters=12, adjoint_mode="forward")
    out = _run(tmp_path)
    m = re.search(r"Vector function derivative is\s+(\S+)", out)
    assert m, out[-2000:]
    dJ = float(m.group(1))
    vx = np.zeros(small_cloud.n)
    vx[0] = 1.0
    with LskumSolver(small_cloud, **CASE) as s:
        J, dJt = s.tangent(12, vx, np.zeros(small_cloud.n))
        _, gx, _, _ = s.adjoint(12, snaps=4)
    assert abs(dJ - dJt) <= 1e-14 * abs(dJt) + 1e-300
    assert abs(dJ - gx[0]) <= 1e-10 * abs(gx[0])


def _d2_sensor(cloud, prim):
    """adaptation_sensors.F90:7-101 restated with numpy: per point the largest D2
    distance to a neighbour's state, over the largest of all points."""
    rho, u1, u2, pr = prim.T
    s = np.zeros(cloud.n)
    for i in range(cloud.n):
        j = cloud.conn[i, :cloud.nbhs[i]]
        t1 = (rho[j] - rho[i]) * np.log((pr[i] * pr[i] * rho[j]) / (pr[j] * pr[j] * rho[i]))
        usq = (u1[j] - u1[i]) ** 2 + (u2[j] - u2[i]) ** 2
        t2 = (0.5 * rho[i] * rho[i] / pr[i]) * ((rho[j] / rho[i]) + (pr[i] * rho[j]) / (pr[j] * rho[i])) * usq
        t3 = 2 * (rho[i] * ((pr[i] * rho[j]) / (pr[j] * rho[i]) - 1) + rho[j] * ((pr[j] * rho[i]) / (pr[i] * rho[j]) - 1))
        s[i] = max(0.0, float((t1 + t2 + t3).max()))
    return s / s.max()


def test_host_input_nml_output_dat(small_cloud, tmp_path):
    """The CUDA build's run: input.nml in, output.dat out (post_processing.F90:7-36):
    x, y, flag_1, qtdepth, prim, the Mach value as that routine computes it, the
    entropy measure and the normalised D2 sensor; then a restart from restart.dat."""
    from mfcfd_amd.solver import LskumSolver
    (tmp_path / "point").mkdir()
    pointio.write_point_h5(str(tmp_path / "point" / "point.h5"), small_cloud)
    nml = ("&input_parameters\n cfl = 0.2\n max_iters = 20\n nsave = 10\n mach = 0.63\n aoa = 2\n"
           " vl_const = 5\n tscheme = 'ssprk43'\n inner_iterations = 2\n blockx = 64\n/\n")
    (tmp_path / "input.nml").write_text(nml)
    _run(tmp_path)
    with open(tmp_path / "output.dat") as fh:
        head = fh.readline().split()
        rows = np.loadtxt(fh)
    with LskumSolver(small_cloud, **CASE) as s:
        s.q_lskum(20)
        prim = s.prim
        s.q_lskum(10, first_it=21)
        prim30 = s.prim
    assert (int(head[0]), int(head[1])) == (small_cloud.n, 20)
    assert np.array_equal(rows[:, 0], small_cloud.x) and np.array_equal(rows[:, 1], small_cloud.y)
    assert np.array_equal(rows[:, 2], small_cloud.flag_1)
    # autotuned launch shapes reorder sums only
    assert rel_err(rows[:, 4:8], prim) <= 1e-11
    rho, u1, u2, pr = rows[:, 4:8].T
    mach_as_written = np.sqrt((u1 * u1 + u2 * u2) / np.sqrt(1.4 * pr / rho))
    assert rel_err(rows[:, 8], mach_as_written) <= 1e-14
    entropy = np.abs(np.log(pr / rho ** 1.4) - np.log(1.0 / 1.4))
    assert np.abs(rows[:, 9] - entropy).max() <= 1e-13
    sensor = _d2_sensor(small_cloud, rows[:, 4:8])
    assert rows[:, 10].max() == 1.0 and np.abs(rows[:, 10] - sensor).max() <= 1e-12
    # restart.dat (initial_conditions.F90:26-37): 10 more iterations continue the run
    shutil.copy(tmp_path / "output.dat", tmp_path / "restart.dat")
    (tmp_path / "input.nml").write_text(nml.replace("max_iters = 20", "max_iters = 10")
                                        .replace("/\n", " restart_solution = 'yes'\n/\n"))
    _run(tmp_path)
    with open(tmp_path / "output.dat") as fh:
        head = fh.readline().split()
        rows = np.loadtxt(fh)
    assert int(head[1]) == 30
    assert rel_err(rows[:, 4:8], prim30) <= 1e-10
