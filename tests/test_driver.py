"""The C++ host driver (congfd): info.txt handling, case set-up, the
stepsLoopCFL loop and its output files."""
import os
import subprocess

import numpy as np
import pytest

from cfd_inwsl_amd import cases
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cfd_inwsl_amd", "congfd")


def _info_txt(path, inter, end_step, output_dt, cfl, ncase, zone, imax, dim, threads=1):
    lines = [inter, end_step, output_dt, cfl, ncase, *zone, *imax, dim, threads]
    with open(path, "w") as f:
        f.write("\n".join(str(x) for x in lines) + "\n")


def _run(args, cwd):
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=600)


def test_driver_rejects_cases_outside_the_path(tmp_path):
    _info_txt(tmp_path / "info.txt", 11, 1, 0.1, 0.5, 3, [0, 0.25, 0, 1, 0, 0], [65, 257, 2], 2)
    r = _run([], tmp_path)
    assert r.returncode == 2
    assert "file mode initialization finished" in r.stdout
    assert "gravity" in r.stderr
    r = _run(["--case", "4"], tmp_path)
    assert r.returncode == 2 and "DoubleMachUp" in r.stderr


def test_driver_usage_error(tmp_path):
    r = _run(["info.txt", "--bogus"], tmp_path)
    assert r.returncode == 2 and "usage" in r.stderr


def _oracle_steps_cfl(grid, cons, cfl, output_dt, end_step):
    """BlockSolver::stepsLoopCFL (zone_FD/src/src/blockSolver.cpp:170-233) over the oracle."""
    t, out_t, step = 0.0, 0.0, 0
    while step < end_step:
        dt = O.cfl_dt(grid, cons, cfl)
        if out_t + dt > output_dt:
            dt = output_dt - out_t
            out_t = 0.0
            step += 1
        else:
            out_t += dt
        cons, _ = O.rk3(grid, cons, dt=dt, nsteps=1)
        t += dt
    return cons, t


def _read_tecplot(path, nvar):
    rows = []
    with open(path) as f:
        for line in f:
            if line[:1].isdigit() or line[:1] == "-":
                rows.append([float(x) for x in line.split()])
    return np.array(rows)


@pytest.mark.gpu
def test_driver_sod_matches_oracle_and_exact(tmp_path):
    """Config S: Sod, 400 cells on [0,1] (reference frame [-0.5,0.5] split at 0),
    WCNS5-JS component-wise + HLLC + SSP-RK3, CFL 0.5, to t = 0.2."""
    _info_txt(tmp_path / "info.txt", 2, 1, 0.2, 0.5, 0, [-0.5, 0.5, 0, 0, 0, 0], [401, 2, 2], 1)
    r = _run(["--recon", "component", "--flux", "hllc"], tmp_path)
    print(r.stdout[-400:], r.stderr[-400:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "Sod - WCNS5 - 401 - t=0.2000.dat"
    assert out.exists()
    log = tmp_path / "Sod - WCNS5 - 401 - rholinf.txt"
    assert log.exists() and log.read_text().count("time = ") > 10
    got = _read_tecplot(out, 4)  # x, rho, u, p
    # oracle through the same loop
    n, h = 400, 1.0 / 400
    x = -0.5 + (np.arange(n) + 0.5) * h
    cons = np.stack([np.where(x < 0, 1.0, 0.125), np.zeros(n),
                     np.where(x < 0, 1.0 / (1.4 - 1) * 1, 1.0 / (1.4 - 1) * 0.1)], -1)
    grid = O.make_grid(1, (n,), h, O.S_JS, O.F_HLLC, 0, [7, 7])
    ref, t = _oracle_steps_cfl(grid, cons, 0.5, 0.2, 1)
    assert abs(t - 0.2) < 1e-14
    assert np.array_equal(got[:, 1], ref[:, 0]), np.abs(got[:, 1] - ref[:, 0]).max()
    exact = cases.sod_exact(x + 0.5, 0.2)
    l1 = np.abs(got[:, 1] - exact[:, 0]).mean()
    print("Sod L1(rho) vs exact:", l1)
    assert l1 < 5e-3
