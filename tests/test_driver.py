"""The C++ host driver (congfd): info.txt handling, case set-up, the
stepsLoopCFL loop and its output files."""
import math
import os
import subprocess

import numpy as np
import pytest

from cfd_inwsl_amd import BC, cases, cgns
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
    _info_txt(tmp_path / "info.txt", 11, 1, 0.1, 0.5, 7, [0, 0.25, 0, 1, 0, 0], [65, 257, 2], 2)
    r = _run([], tmp_path)
    assert r.returncode == 2
    assert "file mode initialization finished" in r.stdout
    assert "not defined" in r.stderr


def test_driver_usage_error(tmp_path):
    r = _run(["info.txt", "--bogus"], tmp_path)
    assert r.returncode == 2 and "usage" in r.stderr


def _oracle_steps_cfl(grid, cons, cfl, output_dt, end_step, nthreads=None):
    """BlockSolver::stepsLoopCFL (zone_FD/src/src/blockSolver.cpp:170-233) over the
    oracle, the time advanced as RK3_SSP + solve do it."""
    t, out_t, step = grid.t, 0.0, 0
    while step < end_step:
        dt = O.cfl_dt(grid, cons, cfl, nthreads)
        if out_t + dt > output_dt:
            dt = output_dt - out_t
            out_t = 0.0
            step += 1
        else:
            out_t += dt
        grid.t = t
        cons, t = O.rk3(grid, cons, dt=dt, nsteps=1, nthreads=nthreads)
    return cons, t


def _last_output(d):
    """The solution file of the latest output time ("... - t=<t>.cgns")."""
    files = list(d.glob("*t=*.cgns"))
    assert files, list(d.iterdir())
    return max(files, key=lambda f: float(f.name.rsplit("t=", 1)[1][:-5]))


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
    r = _run(["--recon", "component", "--flux", "hllc", "--format", "tecplot"], tmp_path)
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


def _case_state(ncase, zone, imax):
    """Initial state of the driver's 2D presets 3 (RTI) and 4 (double Mach),
    as Initializer::solInit writes them (zone_FD/src/src/initializer.cpp:396-468)."""
    nx, ny = imax[0] - 1, imax[1] - 1
    h = (zone[1] - zone[0]) / (imax[0] - 1)
    xc = zone[0] + (np.arange(nx) + 0.5) * h
    yc = zone[0] + (np.arange(ny) + 0.5) * h
    X, Y = np.meshgrid(xc, yc, indexing="xy")
    G = 1.4
    cos = np.vectorize(math.cos)  # libm's cos, as the driver calls it
    if ncase == 3:
        heavy = Y <= 0.5
        r = np.where(heavy, 2.0, 1.0)
        p = np.where(heavy, 2.0 * Y + 1.0, Y + 3.0 / 2.0)
        u = np.zeros_like(r)
        v = -0.025 * np.sqrt(G * p / r) * cos(8.0 * np.pi * X)
    else:
        post = Y >= np.sqrt(3) * (X - 1.0 / 6.0)
        r = np.where(post, 8.0, 1.4)
        u = np.where(post, 8.25 * math.cos(math.pi / 6), 0.0)
        v = np.where(post, -8.25 * math.sin(math.pi / 6), 0.0)
        p = np.where(post, 116.5, 1.0)
    cons = np.stack([r, r * u, r * v, 1.0 / (G - 1) * p + r * (u * u + v * v) / 2], -1).reshape(-1, 4)
    return cons, h


@pytest.mark.gpu
@pytest.mark.parametrize("ncase,extra", [(3, ["--gravity"]), (3, []), (4, [])])
def test_driver_rti_and_double_mach_match_oracle(tmp_path, ncase, extra):
    """The reference's 2D cases 3 (Rayleigh-Taylor; gravity source optional, as
    the reference leaves it off) and 4 (double Mach reflection: Dirichlet
    inflow, the x < 1/6 wall wiring, the time-dependent DoubleMachUp top) through
    the driver's stepsLoopCFL, against the oracle through the same loop."""
    if ncase == 3:
        zone, imax, out_dt = [0, 0.25, 0, 1, 0, 0], [33, 129, 2], 0.02
    else:
        zone, imax, out_dt = [0, 4, 0, 1, 0, 0], [121, 31, 2], 0.004
    _info_txt(tmp_path / "info.txt", 11, 1, out_dt, 0.5, ncase, zone, imax, 2)
    r = _run(["--flux", "hllc"] + extra, tmp_path)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    sol = cgns.read_solution(_last_output(tmp_path))
    got_rho = sol["Density"].T.reshape(-1)  # (i, j) -> x fastest
    cons, h = _case_state(ncase, zone, imax)
    nx, ny = imax[0] - 1, imax[1] - 1
    post = [8.0, 8.25 * math.cos(math.pi / 6), -8.25 * math.sin(math.pi / 6), 116.5, 0.0]
    bc_prim = np.zeros((6, 5))
    if ncase == 3:
        bc = [BC.SYMMETRYX, BC.SYMMETRYX, BC.DIRICHLET, BC.DIRICHLET]
        bc_prim[2, :4] = [2.0, 0, 0, 1.0]
        bc_prim[3, :4] = [1.0, 0, 0, 2.5]
    else:
        bc = [BC.DIRICHLET, BC.OUTFLOW, BC.DMR_WALL, BC.DMR_TOP]
        bc_prim[0] = bc_prim[2] = bc_prim[3] = post
    grid = O.make_grid(2, (nx, ny), h, O.S_TENO5, O.F_HLLC, 1, [int(b) for b in bc], bc_prim, x0=zone[0],
                       gravity=(0.0, 1.0, 0.0) if "--gravity" in extra else (0.0, 0.0, 0.0))
    ref, t = _oracle_steps_cfl(grid, cons, 0.5, out_dt, 1)
    assert abs(t - out_dt) < 1e-12
    assert np.array_equal(got_rho, ref[:, 0]), np.abs(got_rho - ref[:, 0]).max()


GOLD = os.path.join(ROOT, "tests", "golden")


def _reference_fixtures():
    import json
    return (json.load(open(os.path.join(GOLD, "reference_outputs.json"))),
            np.load(os.path.join(GOLD, "reference_outputs.npz")))


@pytest.mark.gpu
@pytest.mark.parametrize("key,inter", [(k, i) for k in ("sod", "shu_osher", "lax", "double_rarefaction",
                                                        "woodward_colella")
                                       for i in (11, 8) if not (k in ("shu_osher", "woodward_colella") and i == 8)])
def test_driver_reproduces_reference_solution_files(tmp_path, key, inter):
    """The reference's own solution files (zone_FD/cases: TENO.cgns = Teno5_Z,
    inter 11; TENOS.cgns = Teno5_CongZ, inter 8; char-wise + HLLC, CFL 0.5):
    the GPU through the driver's stepsLoopCFL writes the same Density,
    XVelocity and Pressure, bit for bit."""
    meta, arr = _reference_fixtures()
    m = meta["cases"][key]
    nout = int(round(m["t_end"] / m["output_dt"]))
    _info_txt(tmp_path / "info.txt", inter, nout, m["output_dt"], m["cfl"], m["case"], m["zone"], m["imax"], 1)
    r = _run(["--flux", "hllc"], tmp_path)
    print(key, r.stdout[-200:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    sol = cgns.read_solution(_last_output(tmp_path))
    tag = key if inter == 11 else key + "_tenos"
    for f in ("Density", "XVelocity", "Pressure"):
        want = arr[f"{tag}/{f}"]
        assert np.array_equal(sol[f], want), (f, np.abs(sol[f] - want).max())


@pytest.mark.gpu
def test_driver_reproduces_the_zyc_research_run(tmp_path):
    """'Sod - whfzycTCNSNmyASF002_1 - 101' (inter 33, TENO-S with C_T
    following eta_min): the driver's t = 0.1 and t = 0.2 files equal the
    reference's, bit for bit.  info.txt takes only values below INTERMAX
    (zoneMain.cpp:261-263; the reference sets research variants in
    zoneMain's source), so the scheme comes from --inter."""
    meta, arr = _reference_fixtures()
    m = meta["sod_zyc"]
    nout = int(round(m["t_end"] / m["output_dt"]))
    _info_txt(tmp_path / "info.txt", m["inter"], nout, m["output_dt"], m["cfl"], m["case"], m["zone"], m["imax"], 1)
    r = _run(["--flux", "hllc", "--inter", str(m["inter"])], tmp_path)
    print(r.stdout[-200:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    for t in ("0.1000", "0.2000"):
        files = sorted(tmp_path.glob(f"*t={t}.cgns"))
        assert len(files) == 1, files
        sol = cgns.read_solution(files[0])
        for f in ("Density", "XVelocity", "Pressure"):
            want = arr[f"sod_zyc/t{t}/{f}"]
            assert np.array_equal(sol[f], want), (t, f, np.abs(sol[f] - want).max())


@pytest.mark.gpu
@pytest.mark.parametrize("key,variant,inter", [(k, "tenoas", i) for k in ("sod", "shu_osher", "lax", "double_rarefaction",
                                                                          "woodward_colella")
                                               for i in (26, 27)] + [(k, "tenolad", 21) for k in ("shu_osher", "lax")])
def test_driver_reproduces_reference_research_files(tmp_path, key, variant, inter):
    """TENOAS.cgns (whf_ai_TcnsN_AS_1 / _AS_2, inter 26 / 27) and
    TENOLAD.cgns (whf_TcnsN_LAD, inter 21) through the driver (--inter, as
    info.txt takes only values below INTERMAX): bit for bit."""
    meta, arr = _reference_fixtures()
    m = meta["cases"][key]
    nout = int(round(m["t_end"] / m["output_dt"]))
    _info_txt(tmp_path / "info.txt", 11, nout, m["output_dt"], m["cfl"], m["case"], m["zone"], m["imax"], 1)
    r = _run(["--flux", "hllc", "--inter", str(inter)], tmp_path)
    print(key, variant, inter, r.stdout[-150:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    sol = cgns.read_solution(_last_output(tmp_path))
    for f in ("Density", "XVelocity", "Pressure"):
        want = arr[f"{key}_{variant}/{f}"]
        assert np.array_equal(sol[f], want), (f, np.abs(sol[f] - want).max())


def _digest(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.gpu
@pytest.mark.parametrize("variant,inter", [("TENO", 11), ("TENOAS", 26)])
def test_driver_reproduces_reference_double_mach_files(tmp_path, variant, inter):
    """Double Mach reflection, 512 x 128 cells, TENO5 (inter 11) and
    whf_ai_TcnsN_AS_1 (inter 26) char-wise + HLLC, CFL 0.5, outputs every 0.01
    to t = 0.2 (zoneMain's Double Mach preset: endStep 20, outputDt 0.01):
    the driver's t = 0.2 file equals the reference's "512, 128 - TENO" /
    "TENOAS" file bit for bit (all four fields).  This run exercises the
    time-dependent DoubleMachUp top, the x < 1/6 wall wiring, HLLC's star
    branches and the 2D positivity fallback at the Mach-10 shock."""
    meta, arr = _reference_fixtures()
    m = meta["dmr"]
    nout = int(round(m["t_end"] / m["output_dt"]))
    _info_txt(tmp_path / "info.txt", 11, nout, m["output_dt"], m["cfl"], m["case"], m["zone"], m["imax"], 2)
    r = _run(["--flux", "hllc", "--inter", str(inter)], tmp_path)
    print(r.stdout[-200:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    sol = cgns.read_solution(_last_output(tmp_path))
    if variant == "TENO":
        d = np.abs(sol["Density"] - arr["dmr/Density_f32"].astype(np.float64))
        print(f"double Mach {variant}: max |drho| vs the file's float32 copy {d.max():.2e}", flush=True)
    for f in ("Density", "XVelocity", "YVelocity", "Pressure"):
        assert _digest(sol[f].T.reshape(-1)) == m["sha256"][variant][f], (variant, f)
