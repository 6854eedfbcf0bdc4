"""Isentropic-vortex convergence on the GPU path (config V): WCNS5-Z
characteristic-wise + Roe + MND6 + SSP-RK3 on the periodic [0,10]^2 box,
error against the exact (translated) solution after one period, with
dt ~ h^(5/3) so that the RK3 error stays below the spatial error."""
import math
import sys

import numpy as np
import pytest

from cfd_inwsl_amd import BC, FluxMethod, InterMethod, Recon, Solver, cases

pytestmark = pytest.mark.gpu


def vortex_error(n, t_end=10.0, base_n=128, base_cfl=0.5):
    cons, h = cases.vortex(n)
    s = Solver(2, (n, n), h, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, [BC.PERIODIC] * 4)
    s.set_state(cons)
    cfl = base_cfl * (base_n / n) ** (2.0 / 3.0)
    dt0 = s.cfl_dt(cfl)
    nsteps = int(math.ceil(t_end / dt0))
    dt = t_end / nsteps
    for _ in range(nsteps):
        s.rk3_step(dt)
    got = s.get_state()
    s.close()
    exact, _ = cases.vortex(n, t=t_end)
    err = np.abs(got[:, 0] - exact[:, 0])
    return err.mean(), err.max(), nsteps


def _orders(errs):
    return [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]


def test_vortex_convergence_order():
    """One period on 128^2 .. 1024^2: the L1 order of density is >= 4.9 on
    both of the two finest pairs (256->512, 512->1024); L1 and Linf orders
    are printed for every pair."""
    ns = [128, 256, 512, 1024]
    res = [vortex_error(n) for n in ns]
    for n, (l1, li, st) in zip(ns, res):
        print(f"vortex n={n} steps={st} L1={l1:.3e} Linf={li:.3e}", flush=True)
    o1 = _orders([r[0] for r in res])
    oi = _orders([r[1] for r in res])
    print("L1 orders", [round(x, 3) for x in o1], "Linf orders", [round(x, 3) for x in oi], flush=True)
    sys.stdout.flush()
    assert o1[-1] >= 4.9 and o1[-2] >= 4.9


def _ten_periods(n, cfl=0.5, periods=10, period=10.0):
    """The benchmark's own stepping (device CFL dt at CFL 0.5, stepsLoopCFL
    clamped at an output every period) for `periods` periods."""
    cons, h = cases.vortex(n)
    s = Solver(2, (n, n), h, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, [BC.PERIODIC] * 4)
    s.set_state(cons)
    out_t, steps = 0.0, 0
    for _ in range(periods):
        reached = False
        while not reached:
            log, out_t, reached = s.steps_to_output(cfl, period, out_t, max_steps=1024)
            steps += len(log)
    t = s.get_time()
    got = s.get_state()
    s.close()
    exact, _ = cases.vortex(n, t=t)
    err = np.abs(got[:, 0] - exact[:, 0])
    # where the largest error sits: distance from the vortex centre (the
    # vortex is back at (5, 5) after whole periods)
    j, i = divmod(int(err.argmax()), n)
    h = 10.0 / n
    print(f"n={n}: Linf at x={(i + 0.5) * h:.3f} y={(j + 0.5) * h:.3f}, "
          f"r={math.hypot((i + 0.5) * h - 5.0, (j + 0.5) * h - 5.0):.3f} from the centre", flush=True)
    return err.mean(), err.max(), steps, t


def test_vortex_ten_periods_at_cfl_half():
    """Ten periods (t = 100) on 512^2 and 1024^2 with the benchmark's CFL 0.5
    and device dt: density L1 / Linf against the exact vortex and their
    orders are printed.  Measured: L1 8.9e-7 -> 2.0e-8 (order 5.5), Linf
    2.8e-5 -> 9.3e-6 (order 1.6; the printed location of the largest error
    says whether it sits at the box edge, where the vortex's 6e-6 tail makes
    its periodic extension kink).  Asserted: L1 order >= 4.5 and L1 at
    1024^2 below 1e-7."""
    res = {n: _ten_periods(n) for n in (512, 1024)}
    for n, (l1, li, st, t) in res.items():
        print(f"vortex 10 periods n={n} steps={st} t={t:.12g} L1={l1:.3e} Linf={li:.3e}", flush=True)
    o1 = math.log2(res[512][0] / res[1024][0])
    oi = math.log2(res[512][1] / res[1024][1])
    print(f"10-period orders 512->1024: L1 {o1:.3f} Linf {oi:.3f}", flush=True)
    assert abs(res[1024][3] - 100.0) < 1e-9
    assert res[1024][0] < 1e-7
    assert o1 >= 4.5
