"""Benchmark: Mcell-stage/s of the CongFD RHS + SSP-RK3 step on one MI355X.

Workload (default): 2D isentropic vortex on a periodic 2048^2 grid, FP64,
WCNS5-Z characteristic-wise + Roe flux + MND6, SSP-RK3 with the CFL time step
computed on the device every step.  A "step" is one full RK3 step (three
stages of consToPrim + ghost fill + x/y sweeps + stage update, plus the CFL
reduction).  Mcell-stage/s = cells * 3 * steps / seconds / 1e6.

The same JSON line carries the 3D 256^3 periodic-flow figure (TENO5
characteristic-wise + HLLC) under "secondary", the live roofline of the
sweep kernel (HIP events on the solver's stream) and a CPU baseline (the
in-repo oracle on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from cfd_inwsl_amd import (BATCH_GRAPH, BATCH_PER_ZONE, BATCH_PERSISTENT, BC, PRECISION_FP32,  # noqa: E402
                           PRECISION_FP64, PRECISION_FP64_FMA, Batch, FluxMethod, InterMethod, Recon, Solver, cases)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(dim):
    """HBM bytes per sweep launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE x 2 + WRITE_SIZE x 1, copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_sweeps.json")
    if not os.path.exists(path):
        return None
    key = "2d_vortex_2048" if dim == 2 else "3d_periodic_256"
    entry = json.load(open(path)).get(key)
    return entry["traffic_bytes_per_launch"] if entry else None


def measure(sol, nv, dim, ncell, steps, warmup, cfl, traffic=True):
    tiles, segs = sol.autotune(reps=3)
    log(f"[bench] launch plan: tiles {tiles} segments {segs}")
    sol.rk3_steps_cfl(cfl, warmup, want_time=False)
    sol.synchronize()
    t0 = time.perf_counter()
    sol.rk3_steps_cfl(cfl, steps, want_time=False)
    sol.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    value = ncell * 3 * steps / sec / 1e6
    # dominant kernel: the direction sweep; algorithmic bytes per cell and
    # direction = 24*nv (direction-split streams, FP64): per-launch bytes
    # = 24*nv*ncell, averaged over the dim sweep launches of a stage
    ms = sol.time_sweeps(reps=20)
    mean_ms = float(np.mean(ms))
    bytes_per_launch = 24 * nv * ncell
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dim) if traffic else None,
            "kernel": "sweep_kernel", "sweep_ms": [round(float(x), 4) for x in ms],
            "bytes_per_launch": bytes_per_launch, "launch_plan": {"tiles": tiles, "segments": segs}}
    return value, sec * 1e3 / steps, roof


def bench_batch(nzones=96, seed=20261203, steps=3):
    """Config B: independent 2D four-quadrant Riemann zones, sides uniform in
    [128, 1024], WCNS5 char-wise + HLLC, outflow; one persistent launch per
    sweep with the RK3 step replayed from a hipGraph vs per-zone launches."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(128, 1025, size=(nzones, 2))
    zones, states = [], []
    for i, (nx, ny) in enumerate(sizes):
        c, h = cases.riemann2d_quadrants(int(nx), int(ny), seed=seed + i)
        zones.append({"n": (int(nx), int(ny)), "h": h, "bc": [BC.OUTFLOW] * 6})
        states.append(c)
    cells = int(sum(int(a) * int(b) for a, b in sizes))
    b = Batch(2, zones, InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR)
    for i, c in enumerate(states):
        b.set_state(i, c)
    b.autotune(reps=1)
    out = {}
    finals = {}
    for name, mode in (("graph", BATCH_GRAPH), ("persistent", BATCH_PERSISTENT), ("per_zone", BATCH_PER_ZONE)):
        for i, c in enumerate(states):
            b.set_state(i, c)
        b.steps_cfl(0.5, 1, mode)
        b.synchronize()
        t0 = time.perf_counter()
        b.steps_cfl(0.5, steps, mode)
        b.synchronize()
        sec = time.perf_counter() - t0
        out[name] = {"ms_per_step": round(sec * 1e3 / steps, 3),
                     "mcell_stage_s": round(cells * 3 * steps / sec / 1e6, 1)}
        log(f"[bench] batch {name}: {out[name]}")
        if mode == BATCH_GRAPH:
            finals[PRECISION_FP64] = [b.get_state(i) for i in range(nzones)]
    b.close()
    # FP32 storage (FP32 sweeps, FP64 state and stage combination): the same
    # 1 + steps graph-replayed steps, timed, and the per-variable difference to
    # the FP64 batch over every zone (relative to max |U_v| over the batch)
    b = Batch(2, zones, InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR, precision=PRECISION_FP32)
    for i, c in enumerate(states):
        b.set_state(i, c)
    b.autotune(reps=1)
    for i, c in enumerate(states):
        b.set_state(i, c)
    b.steps_cfl(0.5, 1, BATCH_GRAPH)
    b.synchronize()
    t0 = time.perf_counter()
    b.steps_cfl(0.5, steps, BATCH_GRAPH)
    b.synchronize()
    sec = time.perf_counter() - t0
    out["graph_fp32"] = {"ms_per_step": round(sec * 1e3 / steps, 3),
                         "mcell_stage_s": round(cells * 3 * steps / sec / 1e6, 1)}
    log(f"[bench] batch graph_fp32: {out['graph_fp32']}")
    f32 = [b.get_state(i) for i in range(nzones)]
    b.close()
    f64 = finals[PRECISION_FP64]
    scale = np.max([np.abs(x).max(axis=0) for x in f64], axis=0)
    l1 = sum(np.abs(a - x).sum(axis=0) for a, x in zip(f32, f64)) / cells / scale
    linf = np.max([np.abs(a - x).max(axis=0) for a, x in zip(f32, f64)], axis=0) / scale
    fp32_err = {"steps": steps + 1, "L1_rel": [float(x) for x in l1], "Linf_rel": [float(x) for x in linf]}
    log(f"[bench] batch fp32 vs fp64: {fp32_err}")
    return {"workload": f"{nzones} independent 2D four-quadrant Riemann zones, sides U[128,1024], "
                        "WCNS5 char-wise + HLLC, FP64 (and FP32 storage: graph_fp32), per-zone CFL 0.5",
            "zones": nzones, "cells": cells, "steps": steps, "modes": out,
            "speedup_graph_vs_per_zone": round(out["per_zone"]["ms_per_step"] / out["graph"]["ms_per_step"], 2),
            "fp32_vs_fp64": fp32_err}


def bench_precisions(steps_err=100):
    """The two opt-in arithmetic modes against the exact FP64 path (equal to
    the oracle bit for bit): FP32 storage + arithmetic in the sweeps (FP64 U^n
    and stage combination), and FP64 with fused multiply-adds contracted.
    Throughput on both headline grids, and the per-variable L1 / Linf
    difference to the FP64 path after `steps_err` RK3 steps of a fixed dt."""
    names = {PRECISION_FP64: "fp64", PRECISION_FP32: "fp32", PRECISION_FP64_FMA: "fp64_fma"}
    out = {}
    for name, dim, n, mk, inter, flux in (
            ("2d_vortex_2048", 2, (2048, 2048), lambda: cases.vortex(2048), InterMethod.WCNSZ5, FluxMethod.ROE),
            ("3d_periodic_256", 3, (256, 256, 256), lambda: cases.periodic3d(256), InterMethod.TCNS5, FluxMethod.HLLC)):
        cons, h = mk()
        ncell = int(np.prod(n))
        res = {}
        finals = {}
        for prec in (PRECISION_FP64, PRECISION_FP32, PRECISION_FP64_FMA):
            s = Solver(dim, n, h, inter, flux, Recon.CHAR, [BC.PERIODIC] * 6, precision=prec)
            s.set_state(cons)
            s.autotune(reps=2)
            dt = s.cfl_dt(0.5)
            s.rk3_step(dt)
            s.synchronize()
            s.set_state(cons)
            t0 = time.perf_counter()
            for _ in range(steps_err):
                s.rk3_step(dt)
            s.synchronize()
            sec = time.perf_counter() - t0
            res[names[prec]] = round(ncell * 3 * steps_err / sec / 1e6, 1)
            finals[prec] = s.get_state()
            s.close()
        b = finals[PRECISION_FP64]
        scale = np.abs(b).max(axis=0)
        diff = {}
        for prec in (PRECISION_FP32, PRECISION_FP64_FMA):
            a = finals[prec]
            diff[names[prec]] = {"L1_rel": [float(x) for x in np.abs(a - b).mean(axis=0) / scale],
                                 "Linf_rel": [float(x) for x in np.abs(a - b).max(axis=0) / scale]}
        out[name] = {"mcell_stage_s": res, "steps": steps_err, "vs_fp64": diff}
        log(f"[bench] precisions {name}: {out[name]}")
    # FP32 storage on the two other single-zone configurations: the Sod tube
    # (400 cells, WCNS5-JS component-wise + HLLC, CFL 0.5 to t = 0.2 through
    # the device output loop) and the adjoint workload's forward map (512^2,
    # WCNS5-JS char-wise + HLLC, Mach-1.5 shock, 400 fixed-dt steps)
    def rel(a, b):
        sc = np.abs(b).max(axis=0)
        return {"L1_rel": [float(x) for x in np.abs(a - b).mean(axis=0) / sc],
                "Linf_rel": [float(x) for x in np.abs(a - b).max(axis=0) / sc]}

    sod = {}
    for prec in (PRECISION_FP64, PRECISION_FP32):
        c, h = cases.sod(400)
        s = Solver(1, (400,), h, InterMethod.WCNS5, FluxMethod.HLLC, Recon.COMPONENT, [BC.OUTFLOW] * 6,
                   precision=prec)
        s.set_state(c)
        ot, reached, nst = 0.0, False, 0
        while not reached:
            lg, ot, reached = s.steps_to_output(0.5, 0.2, ot, max_steps=256)
            nst += len(lg)
        sod[prec] = (s.get_state(), s.get_time(), nst)
        s.close()
    out["1d_sod_400"] = {"t": sod[PRECISION_FP32][1], "steps": sod[PRECISION_FP32][2],
                         "vs_fp64": {"fp32": rel(sod[PRECISION_FP32][0], sod[PRECISION_FP64][0])}}
    log(f"[bench] precisions 1d_sod_400: {out['1d_sod_400']}")
    case = cases.adjoint_shock(512)
    U0 = cases.adjoint_state(case)
    dt = cases.cfl_dt_2d(U0, 512, case["h"])
    adj = {}
    for prec in (PRECISION_FP64, PRECISION_FP32):
        s = Solver(2, (512, 512), (case["h"], case["h"]), InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR,
                   [BC.PERIODIC] * 6, precision=prec)
        s.set_state(U0)
        for _ in range(400):
            s.rk3_step(dt)
        adj[prec] = s.get_state()
        s.close()
    out["2d_adjoint_shock_512"] = {"steps": 400, "vs_fp64": {"fp32": rel(adj[PRECISION_FP32], adj[PRECISION_FP64])}}
    log(f"[bench] precisions 2d_adjoint_shock_512: {out['2d_adjoint_shock_512']}")
    return out


def bench_adjoint(n=512, nsteps=400, every=20, ndir=32):
    """Adjoint workload: 512^2 periodic, WCNS5-JS char-wise + HLLC, Mach-1.5
    shock, 400 fixed-dt RK3 steps, J = int p^2 over a 64^2 window.  Times the
    checkpointed forward and the backward (stage recomputation + three RHS
    vector-Jacobian products per step), and checks dJ/drho0 against central
    differences of the GPU forward and, when present, the dual-number oracle
    values in tests/golden/adjoint_dual_512.json on the same 32 directions."""
    from cfd_inwsl_amd import adjoint
    case = cases.adjoint_shock(n)
    U0 = cases.adjoint_state(case)
    dt = cases.cfl_dt_2d(U0, n, case["h"])
    s = Solver(2, (n, n), (case["h"], case["h"]), InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR,
               [BC.PERIODIC] * 6)
    s.set_state(U0)
    s.adjoint_forward(dt, every, every)
    s.adjoint_backward(np.zeros_like(U0))
    s.set_state(U0)
    t0 = time.perf_counter()
    s.adjoint_forward(dt, nsteps, every)
    t1 = time.perf_counter()
    UT = s.get_state()
    J, gJ = adjoint.window_pressure_objective(UT, case)
    t2 = time.perf_counter()
    gU = s.adjoint_backward(gJ)
    t3 = time.perf_counter()
    g = gU.reshape(n, n, 4)
    u, v = case["u"], case["v"]
    grho = case["rho_factor"] * (g[..., 0] + g[..., 1] * u + g[..., 2] * v + g[..., 3] * 0.5 * (u * u + v * v))
    fwd_ms, bwd_ms = (t1 - t0) * 1e3, (t3 - t2) * 1e3
    log(f"[bench] adjoint: forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms, ratio {bwd_ms / fwd_ms:.2f}")
    # the same with every stage state kept (no recomputation)
    s.set_state(U0)
    t4 = time.perf_counter()
    s.adjoint_forward(dt, nsteps, 0)
    t5 = time.perf_counter()
    _, gJ0 = adjoint.window_pressure_objective(s.get_state(), case)
    t6 = time.perf_counter()
    gU0 = s.adjoint_backward(gJ0)
    t7 = time.perf_counter()
    # (its VJPs re-evaluate the interfaces exactly instead of reading the
    # recording sweeps' records: the same branches, gradients equal to rounding)
    keep_all = {"forward_ms": round((t5 - t4) * 1e3, 2), "backward_ms": round((t7 - t6) * 1e3, 2),
                "adjoint_over_forward": round((t7 - t6) / (t5 - t4), 3),
                "gradient_max_rel_diff": float(np.abs(gU0 - gU).max() / np.abs(gU).max())}
    log(f"[bench] adjoint, every stage kept: {keep_all}")
    # TENO5 (flags from the records), checkpoint every `every`
    st = Solver(2, (n, n), (case["h"], case["h"]), InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR,
                [BC.PERIODIC] * 6)
    st.set_state(U0)
    st.adjoint_forward(dt, every, every)
    st.adjoint_backward(np.zeros_like(U0))
    st.set_state(U0)
    t8 = time.perf_counter()
    st.adjoint_forward(dt, nsteps, every)
    t9 = time.perf_counter()
    _, gJt = adjoint.window_pressure_objective(st.get_state(), case)
    t10 = time.perf_counter()
    st.adjoint_backward(gJt)
    t11 = time.perf_counter()
    st.close()
    teno = {"forward_ms": round((t9 - t8) * 1e3, 2), "backward_ms": round((t11 - t10) * 1e3, 2),
            "adjoint_over_forward": round((t11 - t10) / (t9 - t8), 3)}
    log(f"[bench] adjoint, TENO5: {teno}")

    def Jof(rho):
        s.set_state(cases.adjoint_state(case, rho))
        for _ in range(nsteps):
            s.rk3_step(dt)
        return adjoint.window_pressure_objective(s.get_state(), case)[0]

    rng = np.random.default_rng(99)
    dirs = [cases._random_phase_field_2d(n, 8, rng) for _ in range(ndir)]
    adj = [float((grho * d).sum()) for d in dirs]
    eps = 1e-7
    fd_rel = []
    for d, a in zip(dirs, adj):
        fd = (Jof(case["rho0"] + eps * d) - Jof(case["rho0"] - eps * d)) / (2 * eps)
        fd_rel.append(abs(a - fd) / abs(fd))
    s.close()
    log(f"[bench] adjoint vs central FD (eps {eps}): median {np.median(fd_rel):.2e} max {max(fd_rel):.2e}")
    # the smooth companion workload (acoustic and entropy waves on a Mach-0.58
    # stream: no wave-region or shock-factor switch a perturbation could
    # cross), same grid, steps and checkpoints: the FD agreement of a smooth map
    sc = cases.adjoint_smooth(n)
    Us = cases.adjoint_state(sc)
    dts = cases.cfl_dt_2d(Us, n, sc["h"])
    ss = Solver(2, (n, n), (sc["h"], sc["h"]), InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6)
    ss.set_state(Us)
    ss.adjoint_forward(dts, nsteps, every)
    _, gJs = adjoint.window_pressure_objective(ss.get_state(), sc)
    gs = ss.adjoint_backward(gJs).reshape(n, n, 4)
    us, vs = sc["u"], sc["v"]
    grs = sc["rho_factor"] * (gs[..., 0] + gs[..., 1] * us + gs[..., 2] * vs + gs[..., 3] * 0.5 * (us * us + vs * vs))

    def Jsm(rho):
        ss.set_state(cases.adjoint_state(sc, rho))
        for _ in range(nsteps):
            ss.rk3_step(dts)
        return adjoint.window_pressure_objective(ss.get_state(), sc)[0]

    rng_s = np.random.default_rng(123)
    eps_s = 1e-4
    fd_s = []
    for _ in range(ndir):
        d = cases._random_phase_field_2d(n, 8, rng_s)
        fd = (Jsm(sc["rho0"] + eps_s * d) - Jsm(sc["rho0"] - eps_s * d)) / (2 * eps_s)
        fd_s.append(abs(float((grs * d).sum()) - fd) / abs(fd))
    ss.close()
    log(f"[bench] smooth workload adjoint vs central FD (eps {eps_s}): median {np.median(fd_s):.2e} max {max(fd_s):.2e}")
    out = {"workload": f"{n}^2 periodic, WCNS5-JS char-wise + HLLC, Mach-1.5 shock, {nsteps} fixed-dt RK3 steps, "
                       f"J = int p^2 over 64^2 window, checkpoint every {every}",
           "forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
           "adjoint_over_forward": round(bwd_ms / fwd_ms, 3), "J": J, "keep_every_stage": keep_all,
           "teno5": teno,
           "fd": {"eps": eps, "directions": ndir, "median_rel": float(np.median(fd_rel)), "max_rel": max(fd_rel)},
           "fd_smooth": {"workload": "the same grid, steps and checkpoints on the smooth companion state (acoustic and "
                                     "entropy waves on a Mach-0.58 stream, no shock)",
                         "eps": eps_s, "directions": ndir, "median_rel": float(np.median(fd_s)), "max_rel": max(fd_s)}}
    golden = os.path.join(ROOT, "tests", "golden", "adjoint_dual_512.json")
    if os.path.exists(golden):
        ref = json.load(open(golden))
        k = len(ref["jvp"])
        rel = [abs(a - b) / abs(b) for a, b in zip(adj[:k], ref["jvp"])]
        out["dual_oracle"] = {"directions": k, "max_rel": max(rel), "J_rel": abs(J - ref["J"]) / abs(ref["J"])}
        log(f"[bench] adjoint vs dual oracle ({k} directions): max {max(rel):.2e}")
    return out


def _host_cpu():
    """CPU model string and logical CPU count of the host (Linux)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_threads():
    """Threads for the CPU baseline: the CPUs this process may run on
    (affinity mask), capped by OMP_NUM_THREADS when the host sets it (a
    shared host gives one GPU's job a fixed CPU share that way)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(avail, env) if env > 0 else avail


def cpu_baseline(seconds_target=12.0):
    """In-repo oracle (OpenMP over pencils, like SpDistributor::rhsSolve) timed
    on the host cores, threads pinned one per core, on bounded samples of
    both headline workloads (2D vortex 2048^2 and 3D periodic 256^3)."""
    threads = cpu_threads()
    # pin before the OpenMP runtime starts (the oracle library loads it)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as O
    model, ncpu = _host_cpu()

    def timed(dim, n, mk, sch, flx):
        cons, h = mk(n)
        grid = O.make_grid(dim, (n,) * dim, h, sch, flx, 1, [1] * (2 * dim))
        t0 = time.perf_counter()
        cur, _ = O.rk3(grid, cons, cfl=0.5, nsteps=1, nthreads=threads)
        one = time.perf_counter() - t0
        steps = max(1, min(8, int(seconds_target / max(one, 1e-3)) - 1))
        t0 = time.perf_counter()
        O.rk3(grid, cur, cfl=0.5, nsteps=steps, nthreads=threads)
        sec = time.perf_counter() - t0
        return n ** dim * 3 * steps / sec / 1e6, steps

    v2, st2 = timed(2, 2048, cases.vortex, O.S_Z, O.F_ROE)
    log(f"[bench] CPU 2D: {v2:.2f} Mcell-stage/s")
    v3, st3 = timed(3, 256, cases.periodic3d, O.S_TENO5, O.F_HLLC)
    log(f"[bench] CPU 3D: {v3:.2f} Mcell-stage/s")
    return {"value": round(v2, 3), "unit": "Mcell-stage/s", "cores": threads, "kind": "port",
            "sample": f"vortex 2048^2 WCNS5-Z char + Roe, {st2} CFL RK3 steps after 1 untimed",
            "host_cpu": model, "host_logical_cpus": ncpu,
            "threads_note": "one OpenMP thread per available CPU (affinity mask, capped by OMP_NUM_THREADS), "
                            "OMP_PROC_BIND=close, OMP_PLACES=cores",
            "secondary": {"value": round(v3, 3), "unit": "Mcell-stage/s",
                          "sample": f"periodic 256^3 TENO5 char + HLLC, {st3} CFL RK3 steps after 1 untimed"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D 256^3 secondary figure")
    ap.add_argument("--no-large", action="store_true", help="skip the 2D 4096^2 / 3D 384^3 figures")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-zones figure")
    ap.add_argument("--no-precisions", "--no-fp32", dest="no_precisions", action="store_true",
                    help="skip the FP32-storage and FP64-FMA figures")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint figures")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the path would shard by slabs along the slowest axis, with a 4-layer halo exchange per "
              "stage and a max all-reduce of the CFL speeds per step)", file=sys.stderr)
        sys.exit(2)

    cfl = 0.5
    n = 2048
    head_cells = n * n
    log(f"[bench] 2D vortex {n}^2 setup")
    cons, h = cases.vortex(n)
    sol = Solver(2, (n, n), h, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, [BC.PERIODIC] * 4)
    sol.set_state(cons)
    log(f"[bench] timing {args.steps} steps after {args.warmup} warmup")
    value, ms_step, roof = measure(sol, 4, 2, n * n, args.steps, args.warmup, cfl)
    sol.close()
    log(f"[bench] 2D: {value:.1f} Mcell-stage/s, {ms_step:.3f} ms/step, roofline {roof['frac']:.3f}")

    secondary = None
    if not args.no_3d:
        n3 = 256
        log(f"[bench] 3D periodic {n3}^3 setup")
        c3, h3 = cases.periodic3d(n3)
        s3 = Solver(3, (n3, n3, n3), h3, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6)
        s3.set_state(c3)
        st3 = max(2, args.steps // 4)
        v3, ms3, roof3 = measure(s3, 5, 3, n3 ** 3, st3, max(1, args.warmup // 2), cfl)
        s3.close()
        log(f"[bench] 3D: {v3:.1f} Mcell-stage/s, {ms3:.3f} ms/step, roofline {roof3['frac']:.3f}")
        secondary = {"workload": "3D periodic random-phase flow 256^3, TENO5 char-wise + HLLC, FP64",
                     "value": round(v3, 2), "unit": "Mcell-stage/s", "ms_per_step": round(ms3, 4),
                     "steps": st3, "roofline": roof3}

    large = None
    if not args.no_large:
        large = {}
        nl = 4096
        log(f"[bench] 2D vortex {nl}^2 setup")
        cons, h = cases.vortex(nl)
        sol = Solver(2, (nl, nl), h, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, [BC.PERIODIC] * 4)
        sol.set_state(cons)
        del cons
        st = max(2, args.steps // 4)
        v, ms, rf = measure(sol, 4, 2, nl * nl, st, max(1, args.warmup // 2), cfl, traffic=False)
        sol.close()
        log(f"[bench] 2D {nl}^2: {v:.1f} Mcell-stage/s, {ms:.3f} ms/step, roofline {rf['frac']:.3f}")
        large["vortex2d_4096"] = {"value": round(v, 2), "unit": "Mcell-stage/s", "ms_per_step": round(ms, 4),
                                  "steps": st, "roofline": rf}
        n3 = 384
        log(f"[bench] 3D periodic {n3}^3 setup")
        c3, h3 = cases.periodic3d(n3)
        s3 = Solver(3, (n3, n3, n3), h3, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6)
        s3.set_state(c3)
        del c3
        st3 = max(2, args.steps // 8)
        v3, ms3, rf3 = measure(s3, 5, 3, n3 ** 3, st3, 1, cfl, traffic=False)
        s3.close()
        log(f"[bench] 3D {n3}^3: {v3:.1f} Mcell-stage/s, {ms3:.3f} ms/step, roofline {rf3['frac']:.3f}")
        large["periodic3d_384"] = {"value": round(v3, 2), "unit": "Mcell-stage/s", "ms_per_step": round(ms3, 4),
                                   "steps": st3, "roofline": rf3}

    batch = None
    if not args.no_batch:
        log("[bench] batched zones")
        batch = bench_batch()

    precisions = None
    if not args.no_precisions:
        log("[bench] FP32-storage and FP64-FMA variants")
        precisions = bench_precisions()

    adj = None
    if not args.no_adjoint:
        log("[bench] adjoint")
        adj = bench_adjoint()

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline()
        log(f"[bench] CPU: {cpu['value']} Mcell-stage/s on {cpu['cores']} threads")

    line = {
        "metric": "Mcell-stage/s at 2D 2048² & 3D 256³ (% HBM roofline); adjoint÷forward time",
        "value": round(value, 2),
        "unit": "Mcell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D isentropic vortex 2048^2 periodic, WCNS5-Z char-wise + Roe + MND6, "
                               "SSP-RK3 with device CFL dt (CFL 0.5)",
                   "cells": head_cells, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "secondary": secondary,
        "large": large,
        "batch": batch,
        "precisions": precisions,
        "adjoint": adj,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
