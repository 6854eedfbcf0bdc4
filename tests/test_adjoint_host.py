"""CPU pins of the adjoint's reference points.

* The dual-number build of the oracle: its values equal the plain oracle bit
  for bit, and its tangents converge to central finite differences at second
  order (a wrong derivative rule stalls the convergence).
* The host statement of the RHS vector-Jacobian product (adjoint_check.cpp,
  built from the same point kernels the GPU adjoint runs) against the dual
  oracle through <(dR/dU)^T mu, d> = <mu, (dR/dU) d>, for every boundary
  family, and the register-lean interface pass against the plain one.
* The SSP-RK3 reverse recurrences on the host against the dual RK3.
"""
import ctypes
import os

import numpy as np
import pytest

from cfd_inwsl_amd import BC, FluxMethod, InterMethod, Recon, cases
from oracle import oracle as O
from tests.helpers import oracle_grid, random_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "adjcheck", "libadjcheck.so")
PD = ctypes.POINTER(ctypes.c_double)
PI = ctypes.POINTER(ctypes.c_int)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "adjcheck")], check=True)
    L = ctypes.CDLL(LIB)
    L.adj_rhs_vjp.argtypes = [PI, PD, PI, PD, PD, PD, PD]
    L.adj_rhs_vjp_s.argtypes = [PI, PD, PI, PD, PD, PD, PD, ctypes.c_int, ctypes.c_double]
    L.adj_interface_vjp.argtypes = [ctypes.c_int, PD, PD, PD]
    L.adj_interface_vjp_lean.argtypes = [ctypes.c_int, PD, PD, PD]
    return L


def _p(a):
    return a.ctypes.data_as(PD)


def host_vjp(lib, n, h, bc, bc_prim, U, mu, teno_ct=None):
    """(d rhs / d U)^T mu on the host: WENO-JS, or TENO5 with cut-off teno_ct."""
    nn = (ctypes.c_int * 2)(*n)
    hh = (ctypes.c_double * 2)(*h)
    bb = (ctypes.c_int * 4)(*[int(b) for b in bc])
    bp = np.ascontiguousarray(np.asarray(bc_prim, dtype=np.float64)[:4, :4])
    U = np.ascontiguousarray(U, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    g = np.zeros_like(U)
    if teno_ct is None:
        lib.adj_rhs_vjp(nn, hh, bb, _p(bp), _p(U), _p(mu), _p(g))
    else:
        lib.adj_rhs_vjp_s(nn, hh, bb, _p(bp), _p(U), _p(mu), _p(g), 1, float(teno_ct))
    return g


def _grid(n, h, bc, bc_prim=None):
    return oracle_grid(2, n, h, InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR, list(bc) + [BC.OUTFLOW] * 2,
                       bc_prim)


def test_dual_values_equal_oracle():
    cons, h = cases.vortex(24)
    g = _grid((24, 24), h, [BC.PERIODIC] * 4)
    d = np.random.default_rng(0).normal(size=cons.shape)
    r, _ = O.rhs(g, cons)
    rd, _ = O.dual_rhs(g, cons, d)
    assert np.array_equal(r, rd)
    u, _ = O.rk3(g, cons, dt=0.05, nsteps=3)
    ud, _ = O.dual_rk3(g, cons, d, 0.05, 3)
    assert np.array_equal(u, ud)
    # a 3D TENO configuration goes through the same template
    c3 = random_state(3, (10, 9, 8), seed=4)
    g3 = oracle_grid(3, (10, 9, 8), 0.1, InterMethod.TCNS5, FluxMethod.ROE, Recon.CHAR, [BC.PERIODIC] * 6)
    assert np.array_equal(O.rhs(g3, c3)[0], O.dual_rhs(g3, c3, np.ones_like(c3))[0])


@pytest.mark.parametrize("jumps", [False, True])
def test_dual_tangent_converges_to_central_differences(jumps):
    n = (20, 18)
    cons = random_state(2, n, seed=11, jumps=jumps)
    g = _grid(n, 0.1, [BC.PERIODIC] * 4)
    rng = np.random.default_rng(1)
    d = rng.normal(size=cons.shape) * np.abs(cons)
    _, jd = O.dual_rhs(g, cons, d)
    errs = []
    for eps in (1e-5, 1e-6):
        fd = (O.rhs(g, cons + eps * d)[0] - O.rhs(g, cons - eps * d)[0]) / (2 * eps)
        errs.append(np.abs(fd - jd).max() / np.abs(jd).max())
    # second-order convergence: 10x smaller step, ~100x smaller error
    assert errs[1] < 2e-6
    assert errs[1] < errs[0] / 30


def test_lean_interface_pass_equals_plain(lib):
    rng = np.random.default_rng(5)
    for trial in range(40):
        w = np.empty((6, 4))
        w[:, 0] = rng.uniform(0.3, 2.0, 6)
        w[:, 1:3] = rng.uniform(-1.5, 1.5, (6, 2))
        w[:, 3] = rng.uniform(0.2, 2.0, 6)
        lF = rng.normal(size=4)
        for ax in (0, 1):
            a = np.zeros(24)
            b = np.zeros(24)
            lib.adj_interface_vjp(ax, _p(w), _p(lF), _p(a))
            lib.adj_interface_vjp_lean(ax, _p(w), _p(lF), _p(b))
            np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-13 * np.abs(a).max())


BC_SETS = {
    "periodic": ([BC.PERIODIC] * 4, None),
    "outflow": ([BC.OUTFLOW] * 4, None),
    "symmetry": ([BC.SYMMETRYX, BC.SYMMETRYX, BC.SYMMETRYY, BC.SYMMETRYY], None),
    "mirror_dirichlet": ([BC.MIRROR_COPY, BC.DIRICHLET, BC.DIRICHLET, BC.OUTFLOW],
                         [[0, 0, 0, 0, 0], [1.2, 0.3, -0.1, 0.9, 0], [0.8, -0.2, 0.4, 1.1, 0], [0] * 5,
                          [0] * 5, [0] * 5]),
}


TENO_CT = {InterMethod.WCNS5: None, InterMethod.TCNS5: 1e-5, InterMethod.TCNS5CT4: 1e-4}


@pytest.mark.parametrize("inter", list(TENO_CT))
@pytest.mark.parametrize("name", list(BC_SETS))
def test_host_rhs_vjp_matches_dual_oracle(lib, name, inter):
    bc, bc_prim = BC_SETS[name]
    n = (21, 17)
    h = (0.07, 0.09)
    cons = random_state(2, n, seed=23, jumps=True)
    g = oracle_grid(2, n, h, inter, FluxMethod.HLLC, Recon.CHAR, list(bc) + [BC.OUTFLOW] * 2, bc_prim)
    rng = np.random.default_rng(9)
    for _ in range(4):
        d = rng.normal(size=cons.shape)
        mu = rng.normal(size=cons.shape)
        _, jd = O.dual_rhs(g, cons, d)
        gt = host_vjp(lib, n, h, bc, bc_prim if bc_prim is not None else np.zeros((6, 5)), cons, mu,
                      TENO_CT[inter])
        lhs, rhs = float((gt * d).sum()), float((mu * jd).sum())
        scale = np.abs(gt * d).sum()
        assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)


def host_rk3_adjoint(lib, g, n, h, bc, U0, dt, nsteps, lam):
    """Reverse sweep of nsteps fixed-dt RK3 steps on the host (oracle forward)."""
    states = []
    u = U0.copy()
    for _ in range(nsteps):
        u1 = u - dt * O.rhs(g, u)[0]
        u2 = 0.75 * u - 0.25 * dt * O.rhs(g, u1)[0] + 0.25 * u1
        states.append((u, u1, u2))
        u = 1.0 / 3.0 * u - 2.0 / 3.0 * dt * O.rhs(g, u2)[0] + 2.0 / 3.0 * u2
    zero = np.zeros((6, 5))
    l3 = lam.copy()
    for u0, u1, u2 in reversed(states):
        l2 = 2.0 / 3.0 * l3 - 2.0 / 3.0 * dt * host_vjp(lib, n, h, bc, zero, u2, l3)
        l1 = 0.25 * l2 - 0.25 * dt * host_vjp(lib, n, h, bc, zero, u1, l2)
        l3 = 1.0 / 3.0 * l3 + 0.75 * l2 + l1 - dt * host_vjp(lib, n, h, bc, zero, u0, l1)
    return u, l3


def test_host_rk3_reverse_matches_dual_rk3(lib):
    # A generic state (no pressure equilibrium, hence no tied HLLC predicates:
    # where a predicate is tied the two implementations' rounding may take
    # different one-sided derivatives).
    n = (16, 14)
    h = (0.08, 0.08)
    U0 = random_state(2, n, seed=31, jumps=True)
    bc = [BC.PERIODIC] * 4
    g = _grid(n, h[0], bc)
    dt = 0.2 * cases.cfl_dt_2d(U0, 16, h[0])
    rng = np.random.default_rng(2)
    lam = rng.normal(size=U0.shape)
    uT, grad = host_rk3_adjoint(lib, g, n, h, bc, U0, dt, 4, lam)
    for _ in range(3):
        d = rng.normal(size=U0.shape)
        uTd, jd = O.dual_rk3(g, U0, d, dt, 4)
        assert np.array_equal(uTd, O.rk3(g, U0, dt=dt, nsteps=4)[0])
        lhs, rhs = float((grad * d).sum()), float((lam * jd).sum())
        assert abs(lhs - rhs) <= 1e-11 * np.abs(grad * d).sum(), (lhs, rhs)


def test_window_objective_gradient():
    from cfd_inwsl_amd import adjoint
    case = cases.adjoint_shock(32)
    U = cases.adjoint_state(case)
    J, gJ = adjoint.window_pressure_objective(U, case)
    d = np.random.default_rng(4).normal(size=U.shape)
    eps = 1e-6
    fd = (adjoint.window_pressure_objective(U + eps * d, case)[0]
          - adjoint.window_pressure_objective(U - eps * d, case)[0]) / (2 * eps)
    assert abs(fd - (gJ * d).sum()) < 1e-7 * abs(fd)
