"""Runs one workload for a few RK3 steps (or a plain copy) so that rocprofv3
can attribute kernel time and hardware counters; prints one line per step
group.

    python -m cfd_inwsl_amd.profile_run --case vortex2d --n 2048 --steps 5
    python -m cfd_inwsl_amd.profile_run --case periodic3d --n 256 --steps 2
    python -m cfd_inwsl_amd.profile_run --case vortex2d --n 2048 --copy 5
    python -m cfd_inwsl_amd.profile_run --case adjoint2d --n 512 --steps 40
    python -m cfd_inwsl_amd.profile_run --summarize <rocprofv3 output dir>/run_results.db [...]

--autotune picks the launch plan by timing first (as bench.py does); the plan
kept is printed, so its dispatches can be told apart by grid size.
--summarize prints a markdown table per kernel and grid of the rocprofv3
database(s): mean duration (kernel trace) or mean counter value per dispatch
(--pmc runs).
"""
import argparse
import sqlite3
import sys
import time

from cfd_inwsl_amd import BC, FluxMethod, InterMethod, Recon, Solver, cases


def make(case, n):
    if case == "vortex2d":
        cons, h = cases.vortex(n)
        return Solver(2, (n, n), h, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, [BC.PERIODIC] * 4), cons
    if case == "adjoint2d":
        c = cases.adjoint_shock(n)
        return (Solver(2, (n, n), (c["h"], c["h"]), InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR,
                       [BC.PERIODIC] * 4), cases.adjoint_state(c))
    if case == "periodic3d":
        cons, h = cases.periodic3d(n)
        return Solver(3, (n, n, n), h, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6), cons
    raise SystemExit(f"unknown case {case}")


def summarize(paths):
    for path in paths:
        con = sqlite3.connect(path)
        cur = con.cursor()
        views = {r[0] for r in cur.execute("select name from sqlite_master where type='view'")}
        print(f"### {path}\n")
        rows = []
        if "counters_collection" in views:
            rows = list(cur.execute(
                "select kernel_name, grid_size_x, grid_size_y, counter_name, count(*), avg(value) "
                "from counters_collection group by kernel_name, grid_size_x, grid_size_y, counter_name "
                "order by avg(value) desc"))
        if rows:
            print("| kernel | grid x | grid y | counter | dispatches | mean per dispatch |")
            print("|---|---|---|---|---|---|")
            for k, gx, gy, name, n, v in rows:
                print(f"| `{k[:90]}` | {gx} | {gy} | {name} | {n} | {v:.1f} |")
        else:
            print("| kernel | grid x | grid y | dispatches | mean us | min us | max us |")
            print("|---|---|---|---|---|---|---|")
            for k, gx, gy, n, a, lo, hi in cur.execute(
                    "select name, grid_x, grid_y, count(*), avg(duration), min(duration), max(duration) "
                    "from kernels group by name, grid_x, grid_y order by sum(duration) desc"):
                print(f"| `{k[:90]}` | {gx} | {gy} | {n} | {a / 1e3:.1f} | {lo / 1e3:.1f} | {hi / 1e3:.1f} |")
        print()
        con.close()


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize":
        summarize(sys.argv[2:])
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="vortex2d")
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--copy", type=int, default=0, help="run the plain copy probe this many times instead")
    ap.add_argument("--autotune", action="store_true", help="time the launch candidates first")
    a = ap.parse_args()
    s, cons = make(a.case, a.n)
    s.set_state(cons)
    if a.autotune:
        tiles, segs = s.autotune(reps=3)
        print(f"[profile_run] autotuned plan: tiles {tiles} segments {segs}", flush=True)
    print(f"[profile_run] launch plan: {s.launch_plan()}", flush=True)
    print(f"[profile_run] {a.case} n={a.n} ready", flush=True)
    t0 = time.perf_counter()
    if a.copy:
        s.copy_probe(a.copy)
        nbytes = 2 * s.ncell * s.nv * 8
        print(f"[profile_run] copy probe x{a.copy}: {nbytes} bytes moved per launch", flush=True)
    elif a.case == "adjoint2d":
        from cfd_inwsl_amd import adjoint
        c = cases.adjoint_shock(a.n)
        dt = cases.cfl_dt_2d(cons, a.n, c["h"])
        s.adjoint_forward(dt, a.steps, 20)
        t1 = time.perf_counter()
        _, gJ = adjoint.window_pressure_objective(s.get_state(), c)
        t2 = time.perf_counter()
        s.adjoint_backward(gJ)
        t3 = time.perf_counter()
        print(f"[profile_run] adjoint {a.steps} steps: forward {t1 - t0:.3f} s, backward {t3 - t2:.3f} s",
              flush=True)
    else:
        s.rk3_steps_cfl(0.5, a.steps, want_time=False)
        s.synchronize()
        print(f"[profile_run] {a.steps} steps in {time.perf_counter() - t0:.3f} s", flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
