"""The opt-in FMA-contracted FP64 mode (CFD_PRECISION_FP64_FMA) against the
CPU oracle (the reference arithmetic, operation by operation).

Contraction changes the rounding of each interface flux by an ulp or two
and the flux derivative divides that by h, so the difference is measured on
the scale the right-hand side has at grid resolution, R_v = S_v * lam / h
(S_v = max|U_v| for density and energy, max rho*(|u| + c) for every momentum
component, lam = max(|u| + c)): a per-variable max|rhs_v| is no scale where
rhs_v is rounding noise (v-momentum of the x-only shock workload: max|rhs|
1.6e-12) or small beside the flux (the 3D flow's momentum rhs, 0.01 against
a pressure flux of 8).  Bar: ||d_v||_inf / R_v <= 1e-12 for one right-hand
side, and ||d_v||_inf / S_v <= 1e-12 for the state after ten CFL RK3 steps
(against the exact path, equal to the oracle), at the configurations' sizes:
S (1D Sod 400, WENO-JS component-wise + HLLC), V (2D vortex 2048^2, WENO-Z
char + Roe), T (3D periodic 256^3, TENO5 char + HLLC), A (2D 512^2 shock
workload, WENO-JS char + HLLC).  A branch predicate tied at rounding level
(a TENO cut-off, an HLLC wave region, the positivity fallback) may be decided
the other way under contraction; a cell above the bar is counted as such a
flip, and the count is reported and bounded (at most 1e-4 of the cells).  The
per-variable max|rhs_v| difference is printed too."""
import sys

import numpy as np
import pytest

from cfd_inwsl_amd import (BC, PRECISION_FP64, PRECISION_FP64_FMA, FluxMethod, InterMethod, Recon, Solver,
                           cases)
from oracle import oracle as O
from tests.helpers import oracle_grid

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scales(dim, cons):
    """S_v and lam of the module docstring."""
    prim = cases.cons_to_prim(cons)
    rho, p = prim[:, 0], prim[:, -1]
    speed = np.sqrt((prim[:, 1:1 + dim] ** 2).sum(axis=1)) + np.sqrt(cases.GAMMA * p / rho)
    s = np.abs(cons).max(axis=0)
    s[1:1 + dim] = (rho * speed).max()
    return s, float(speed.max())


def _errs(a, b, scale):
    """(max scaled difference outside flipped cells, flipped cells)."""
    d = (np.abs(a - b) / scale).max(axis=1)
    flip = d > TOL
    return (float(d[~flip].max()) if (~flip).any() else 0.0), int(np.count_nonzero(flip))


def _per_variable(a, b):
    m = np.abs(b).max(axis=0)
    return float((np.abs(a - b).max(axis=0) / np.where(m > 0, m, 1.0)).max())


def _case(name):
    if name == "S":
        cons, h = cases.sod(400)
        return 1, (400,), h, cons, InterMethod.WCNS5, FluxMethod.HLLC, Recon.COMPONENT, [BC.OUTFLOW] * 2
    if name == "V":
        cons, h = cases.vortex(2048)
        return 2, (2048, 2048), h, cons, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, [BC.PERIODIC] * 4
    if name == "T":
        cons, h = cases.periodic3d(256)
        return 3, (256, 256, 256), h, cons, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6
    c = cases.adjoint_shock(512)
    return (2, (512, 512), c["h"], cases.adjoint_state(c), InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR,
            [BC.PERIODIC] * 4)


@pytest.mark.parametrize("name", ["S", "V", "T", "A"])
def test_fma_rhs_against_oracle(name):
    dim, n, h, cons, inter, flux, recon, bc = _case(name)
    s = Solver(dim, n, h, inter, flux, recon, bc, precision=PRECISION_FP64_FMA)
    s.set_state(cons)
    got = s.rhs()
    s.close()
    ref = O.rhs(oracle_grid(dim, n, h, inter, flux, recon, bc), cons)[0]
    sv, lam = _scales(dim, cons)
    err, flips = _errs(got, ref, sv * lam / min(np.atleast_1d(h)))
    _say(f"FMA rhs, config {name} {n}: max |d|/R {err:.2e} outside {flips} flipped cells of {len(cons)}; "
         f"per-variable max|d|/max|rhs| {_per_variable(got, ref):.2e}")
    assert flips <= 1e-4 * len(cons) and err <= TOL


@pytest.mark.parametrize("name", ["S", "V", "T", "A"])
def test_fma_steps_against_exact_path(name):
    """Ten CFL RK3 steps (each step's dt from the device): the FMA path's
    state stays within 1e-12 of the exact FP64 path (equal to the oracle)."""
    dim, n, h, cons, inter, flux, recon, bc = _case(name)
    out = {}
    for prec in (PRECISION_FP64, PRECISION_FP64_FMA):
        s = Solver(dim, n, h, inter, flux, recon, bc, precision=prec)
        s.set_state(cons)
        out[prec] = (s.rk3_steps_cfl(0.5, 10), s.get_state())
        s.close()
    a, b = out[PRECISION_FP64_FMA][1], out[PRECISION_FP64][1]
    err, flips = _errs(a, b, _scales(dim, b)[0])
    dt_rel = abs(out[PRECISION_FP64_FMA][0] - out[PRECISION_FP64][0]) / out[PRECISION_FP64][0]
    _say(f"FMA vs exact, config {name}, 10 CFL steps: max |d|/S {err:.2e}, flip cells {flips}, "
         f"time rel {dt_rel:.1e}; per-variable max|d|/max|U| {_per_variable(a, b):.2e}")
    assert flips == 0 and err <= TOL


def test_fma_refusals():
    with pytest.raises(Exception, match="FP64_FMA is built for"):
        Solver(2, (16, 16), 0.1, InterMethod.NEW_WHFTCNSA, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 4,
               precision=PRECISION_FP64_FMA)
    s = Solver(2, (32, 32), 0.1, InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 4,
               precision=PRECISION_FP64_FMA)
    s.set_state(cases.vortex(32)[0])
    with pytest.raises(Exception, match="exact FP64 arithmetic"):
        s.rhs_vjp(np.zeros((32 * 32, 4)))
    s.close()
