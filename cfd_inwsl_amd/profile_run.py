"""Runs one workload for a few RK3 steps (or a plain copy) so that rocprofv3
can attribute kernel time and hardware counters; prints one line per step
group.

    python -m cfd_inwsl_amd.profile_run --case vortex2d --n 2048 --steps 5
    python -m cfd_inwsl_amd.profile_run --case periodic3d --n 256 --steps 2
    python -m cfd_inwsl_amd.profile_run --case vortex2d --n 2048 --copy 5
    python -m cfd_inwsl_amd.profile_run --case adjoint2d --n 512 --steps 40
"""
import argparse
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="vortex2d")
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--copy", type=int, default=0, help="run the plain copy probe this many times instead")
    a = ap.parse_args()
    s, cons = make(a.case, a.n)
    s.set_state(cons)
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
