"""cfd_rk3_steps_cfl_noalloc, the entry point that allocates nothing, through
the C ABI: two calls in a row on one GPU (the second on a smaller grid, so
the scratch array keeps the first call's contents) against
cfd_rk3_steps_cfl on identical handles -- states and clocks equal bit for
bit; a launch plan beyond the scratch capacity is refused."""
import ctypes

import numpy as np
import pytest

from cfd_inwsl_amd import BC, FluxMethod, InterMethod, Recon, Solver, cases
from tests.helpers import random_state

pytestmark = pytest.mark.gpu


def _pair(dim, n, h, inter, flux, bc, cons):
    out = []
    for _ in range(2):
        s = Solver(dim, n, h, inter, flux, Recon.CHAR, bc)
        s.set_state(cons)
        out.append(s)
    return out


def _compare(dim, n, h, inter, flux, bc, cons, steps, label):
    ref, new = _pair(dim, n, h, inter, flux, bc, cons)
    t_ref = ref.rk3_steps_cfl(0.5, steps)
    rc = new.lib.cfd_rk3_steps_cfl_noalloc(new.h, ctypes.c_double(0.5), ctypes.c_int(steps))
    assert rc == 0, new.lib.cfd_last_error().decode()
    new.synchronize()
    a, b = ref.get_state(), new.get_state()
    print(f"{label}: bitwise={np.array_equal(a, b)} t {t_ref} vs {new.get_time()}", flush=True)
    assert np.array_equal(a, b), label
    assert new.get_time() == ref.get_time()
    ref.close()
    new.close()


def test_noalloc_steps_equal_the_allocating_entry_point_twice():
    n = 384
    cons, h = cases.vortex(n)
    _compare(2, (n, n), h, InterMethod.WCNSZ5, FluxMethod.ROE, [BC.PERIODIC] * 4, cons, 4, "vortex 384^2, 4 steps")
    n3 = (40, 36, 30)
    cons3 = random_state(3, n3, seed=21)
    _compare(3, n3, 0.05, InterMethod.TCNS5, FluxMethod.HLLC, [BC.OUTFLOW] * 6, cons3, 3,
             "3D 40x36x30 outflow, 3 steps")


def test_noalloc_refuses_a_plan_beyond_its_scratch():
    n = (64, 64, 64)
    cons = random_state(3, n, seed=3)
    s = Solver(3, n, 0.05, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6)
    s.set_state(cons)
    s.set_launch_plan([0, 1, 1], [64, 64, 1])  # z segments of one cell: 256 x 64 workgroups x 3 maxima
    rc = s.lib.cfd_rk3_steps_cfl_noalloc(s.h, ctypes.c_double(0.5), ctypes.c_int(1))
    assert rc == -1
    assert "scratch capacity" in s.lib.cfd_last_error().decode()
    s.close()
