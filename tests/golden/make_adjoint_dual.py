"""Writes adjoint_dual_512.json: directional derivatives of the adjoint
workload's objective from the dual-number build of the CPU oracle.

Workload: cases.adjoint_shock(512) (periodic unit square, rho0 = 1 + 1e-2 f
with seeded modes |k| <= 8, stationary Mach-1.5 normal shock), characteristic
-wise WCNS5 (WENO-JS) + HLLC, fixed dt = CFL 0.5 on the initial state, 400
SSP-RK3 steps, J = integral of p^2 over the 64 x 64 window behind the shock.
For each of 32 seeded density directions d (cases._random_phase_field_2d with
numpy default_rng(99), in order) the file holds dJ/drho0 . d, exact for the
branches the forward takes (forward-mode duals through the oracle).  The GPU
adjoint gradient dotted with the same directions must reproduce them.

Run from the repository root: python tests/golden/make_adjoint_dual.py [threads]
(about two minutes per direction on 8 cores).
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cfd_inwsl_amd import BC, FluxMethod, InterMethod, Recon, adjoint, cases  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.helpers import oracle_grid  # noqa: E402

N, NSTEPS, NDIR, SEED = 512, 400, 32, 99


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else O.nthreads_default()
    case = cases.adjoint_shock(N)
    U0 = cases.adjoint_state(case)
    dt = cases.cfl_dt_2d(U0, N, case["h"])
    g = oracle_grid(2, (N, N), case["h"], InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 4)
    rng = np.random.default_rng(SEED)
    u, v = case["u"], case["v"]
    out = {"n": N, "nsteps": NSTEPS, "dt": dt, "direction_seed": SEED, "jvp": []}
    for k in range(NDIR):
        d = cases._random_phase_field_2d(N, 8, rng)
        f = case["rho_factor"] * d
        dU = np.stack([f, f * u, f * v, 0.5 * f * (u * u + v * v)], -1).reshape(-1, 4)
        t = time.time()
        uT, duT = O.dual_rk3(g, U0, dU, dt, NSTEPS, nthreads=threads)
        J, gJ = adjoint.window_pressure_objective(uT, case)
        out["J"] = J
        out["jvp"].append(float((gJ * duT).sum()))
        print(f"direction {k}: {out['jvp'][-1]:.15e} ({time.time() - t:.0f} s)", flush=True)
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_dual_512.json"), "w") as fh:
            json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
