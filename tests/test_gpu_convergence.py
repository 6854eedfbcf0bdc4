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
