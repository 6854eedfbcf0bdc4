"""GPU parity against the CPU oracle, through the C ABI.

Bar: FP64 results equal to the oracle.  The kernels keep the reference's
operation order and are compiled without FMA contraction, so the check is
bit-for-bit (np.array_equal); the stated tolerance of the hot path is
||gpu - oracle||_inf / ||oracle||_inf <= 1e-12 per variable, asserted as well
so that a failure reports its size.
"""
import itertools
import sys

import numpy as np
import pytest

from cfd_inwsl_amd import BC, FluxMethod, InterMethod, Recon, Solver, cases
from oracle import oracle as O
from tests.helpers import oracle_grid, random_state, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(gpu, ref, what):
    err = rel_err(gpu, ref)
    _say(f"{what}: rel err {err:.3e} bitwise={np.array_equal(gpu, ref)}")
    assert err <= TOL, what
    assert np.array_equal(gpu, ref), what


SIZES = {1: (61,), 2: (37, 29), 3: (13, 11, 9)}
COMBOS = list(itertools.product(
    [1, 2, 3],
    [InterMethod.WCNS5, InterMethod.WCNSZ5, InterMethod.TCNS5, InterMethod.WCNS5CONGZ],
    [FluxMethod.HLLC, FluxMethod.ROE],
    [Recon.CHAR, Recon.COMPONENT],
))


RESEARCH = [InterMethod.WCNS5CONGZCT4, InterMethod.WCNS5CONGZCT7, InterMethod.CONGTCNS5CT5,
            InterMethod.CONGTCNS5CT10, InterMethod.NEW_WHFTCNSA, InterMethod.WHFZYCTCNSNMYASF002_1,
            InterMethod.NEW_WHFTCNSASF203_NOS, InterMethod.NEW_WHFTCNSASF202_NOS,
            InterMethod.WHFTCNSN, InterMethod.WHFTCNSNA, InterMethod.WHFTCNSNAS, InterMethod.WHFTCNSNS,
            InterMethod.WHFTCNSNLAD, InterMethod.WHFAITCNSNS, InterMethod.WHFAITCNSNA, InterMethod.WHFAITCNSNAS_1,
            InterMethod.WHFAITCNSNAS_2, InterMethod.WHFAITCNSNLADS, InterMethod.WHFAITCNSNAZS,
            InterMethod.WHFAITCNSNMYASF002_1, InterMethod.WHFAITCNSNMYASF002_2, InterMethod.WHFAITCNSNMYASF002_AI1,
            InterMethod.WHFCOMPARE, InterMethod.MUCSLIN5]


@pytest.mark.parametrize("dim,inter", [(d, i) for d in (1, 2, 3) for i in RESEARCH])
def test_rhs_parity_research_variants(dim, inter):
    """The research TENO variants (runtime cut-off constants of TENO5 / TENO-S,
    TENO-A's kernels, and the kernels of TENO-S with C_T following eta_min,
    one runtime law each) against the oracle, characteristic-wise HLLC and
    component-wise Roe, FP64."""
    n = SIZES[dim]
    cons = random_state(dim, n, seed=29, jumps=True)
    for flux, recon in ((FluxMethod.HLLC, Recon.CHAR), (FluxMethod.ROE, Recon.COMPONENT)):
        s = Solver(dim, n, 0.05, inter, flux, recon, [BC.PERIODIC] * 6)
        s.set_state(cons)
        g = s.rhs()
        s.close()
        ref, _ = O.rhs(oracle_grid(dim, n, 0.05, inter, flux, recon, [BC.PERIODIC] * 6), cons)
        what = f"research rhs dim={dim} inter={inter.name} {flux.name} {recon.name}"
        if np.isnan(ref).any():
            # whf_TcnsN_AS keeps the linear stencil across these jumps in
            # component-wise form, and the Roe flux of the negative pressure
            # it interpolates is NaN -- in the oracle as on the GPU
            _say(f"{what}: NaN in {np.isnan(ref).mean():.3f} of the oracle's entries")
            assert np.array_equal(np.isnan(g), np.isnan(ref)), what
            g, ref = np.where(np.isnan(g), 0.0, g), np.where(np.isnan(ref), 0.0, ref)
        _check(g, ref, what)


def test_whf_kernels_are_fp64_single_zone_only():
    """The whf_TcnsN-type schemes (chi / Teno5_Z / whf-S cut-offs) have FP64
    single-zone kernels only: FP32 handles and batches refuse them."""
    with pytest.raises(Exception, match="FP64 only"):
        Solver(2, (40, 40), 0.05, InterMethod.WHFTCNSNLAD, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 4,
               precision=1)


def test_wcns5congzct7_is_teno5_congaa_in_1d():
    """The reference's 1D Euler path wires WCNS5CONGZCT7 to Teno5_CongAA, its
    2D path to Teno5_CongZCT7: the 1D handle equals the oracle's CongAA and
    differs from CT7 on this data."""
    cons = random_state(1, (61,), seed=29, jumps=True)
    s = Solver(1, (61,), 0.05, InterMethod.WCNS5CONGZCT7, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 2)
    s.set_state(cons)
    g = s.rhs()
    s.close()
    ref, _ = O.rhs(O.make_grid(1, (61,), 0.05, O.S_CONGAA, O.F_HLLC, 1, [1, 1]), cons)
    ct7, _ = O.rhs(O.make_grid(1, (61,), 0.05, O.S_TENOS_CT7, O.F_HLLC, 1, [1, 1]), cons)
    assert np.array_equal(g, ref)
    assert not np.array_equal(g, ct7)


@pytest.mark.parametrize("dim,inter,flux,recon", COMBOS)
def test_rhs_parity(dim, inter, flux, recon):
    n = SIZES[dim]
    for bcname, bcv in (("periodic", BC.PERIODIC), ("outflow", BC.OUTFLOW)):
        bc = [bcv] * 6
        cons = random_state(dim, n, seed=dim * 100 + int(inter), jumps=True)
        s = Solver(dim, n, 0.07, inter, flux, recon, bc)
        s.set_state(cons)
        g = s.rhs()
        s.close()
        ref, _ = O.rhs(oracle_grid(dim, n, 0.07, inter, flux, recon, bc), cons)
        _check(g, ref, f"rhs dim={dim} {inter.name} {flux.name} {recon.name} {bcname}")


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_rhs_parity_mixed_boundaries(dim):
    n = {1: (40,), 2: (33, 21), 3: (12, 10, 14)}[dim]
    bc = [BC.SYMMETRYX, BC.DIRICHLET, BC.SYMMETRYY, BC.OUTFLOW, BC.SYMMETRYZ, BC.MIRROR_COPY]
    if dim == 1:
        bc = [BC.SYMMETRY1D, BC.DIRICHLET] + bc[2:]
    bcp = [[0.0] * 5 for _ in range(6)]
    bcp[1] = [0.8, 0.1, -0.2, 0.05, 0.9][: dim + 2] + [0.0] * (3 - dim)
    bcp[1][dim + 1] = 0.9
    cons = random_state(dim, n, seed=7 + dim)
    for inter, flux in ((InterMethod.TCNS5, FluxMethod.HLLC), (InterMethod.WCNS5, FluxMethod.ROE)):
        s = Solver(dim, n, 0.05, inter, flux, Recon.CHAR, bc, bcp)
        s.set_state(cons)
        g = s.rhs()
        s.close()
        ref, _ = O.rhs(oracle_grid(dim, n, 0.05, inter, flux, Recon.CHAR, bc, bcp), cons)
        _check(g, ref, f"rhs mixed-bc dim={dim} {inter.name} {flux.name}")


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_rhs_parity_ct_variants(dim):
    n = SIZES[dim]
    cons = random_state(dim, n, seed=31)
    for inter in (InterMethod.TCNS5CT4, InterMethod.TCNS5CT7):
        s = Solver(dim, n, 0.07, inter, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6)
        s.set_state(cons)
        g = s.rhs()
        s.close()
        ref, _ = O.rhs(oracle_grid(dim, n, 0.07, inter, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6), cons)
        _check(g, ref, f"rhs dim={dim} {inter.name}")


@pytest.mark.parametrize("dim,inter,flux", [(1, InterMethod.WCNS5, FluxMethod.HLLC),
                                            (2, InterMethod.WCNSZ5, FluxMethod.ROE),
                                            (3, InterMethod.TCNS5, FluxMethod.HLLC)])
def test_rk3_parity(dim, inter, flux):
    n = {1: (80,), 2: (40, 36), 3: (16, 12, 10)}[dim]
    bc = [BC.PERIODIC] * 6
    cons = random_state(dim, n, seed=3, jumps=False)
    s = Solver(dim, n, 0.05, inter, flux, Recon.CHAR, bc)
    s.set_state(cons)
    grid = oracle_grid(dim, n, 0.05, inter, flux, Recon.CHAR, bc)
    dt = 0.2 * O.cfl_dt(grid, cons, 1.0)
    for _ in range(3):
        s.rk3_step(dt)
    g = s.get_state()
    ref, _ = O.rk3(grid, cons, dt=dt, nsteps=3)
    _check(g, ref, f"rk3 dim={dim}")
    # CFL-driven steps with the time step computed on the device
    t = s.rk3_steps_cfl(0.5, 2)
    ref2, tref = O.rk3(grid, ref, cfl=0.5, nsteps=2)
    _check(s.get_state(), ref2, f"rk3 cfl dim={dim}")
    assert t == tref
    s.close()


def _dmr_state(n, h, x0):
    nx, ny = n
    x = x0 + (np.arange(nx) + 0.5) * h
    y = x0 + (np.arange(ny) + 0.5) * h
    X, Y = np.meshgrid(x, y, indexing="xy")
    post = [8.0, 8.25 * np.cos(np.pi / 6), -8.25 * np.sin(np.pi / 6), 116.5]
    isp = Y >= np.sqrt(3) * (X - 1.0 / 6.0)
    prim = np.stack([np.where(isp, post[v], [1.4, 0.0, 0.0, 1.0][v]) for v in range(4)], -1)
    return cases.prim_to_cons(prim).reshape(-1, 4), post


def _stepped_parity(s, grid, cons, t0, what):
    """Three fixed-dt steps and two device-CFL steps from absolute time t0,
    against the oracle from the same time; the device clock must read the
    oracle's time, advanced as the reference's RK3 step advances it."""
    s.set_state(cons)
    s.set_time(t0)
    grid.t = t0
    dt = 0.5 * O.cfl_dt(grid, cons, 1.0)
    assert s.cfl_dt(1.0) == 2 * dt and s.get_time() == t0  # a dt query leaves the clock alone
    for _ in range(3):
        s.rk3_step(dt)
    ref, t1 = O.rk3(grid, cons, dt=dt, nsteps=3)
    _check(s.get_state(), ref, what)
    assert s.get_time() == t1, (s.get_time(), t1)
    s.rk3_steps_cfl(0.5, 2)
    grid.t = t1
    ref2, t2 = O.rk3(grid, ref, cfl=0.5, nsteps=2)
    _check(s.get_state(), ref2, what + " cfl")
    assert s.get_time() == t2
    s.close()


@pytest.mark.parametrize("flux", [FluxMethod.HLLC, FluxMethod.ROE])
def test_rk3_parity_double_mach(flux):
    """Double Mach reflection boundaries: the post-shock Dirichlet inflow, the
    wall (Dirichlet for x < 1/6, reflecting beyond; cells placed from the
    zone origin) and the DoubleMachUp top, whose ghost states follow the
    shock with the stage time."""
    n, h, x0 = (90, 24), 1.0 / 24, -0.05
    cons, post = _dmr_state(n, h, x0)
    bc = [BC.DIRICHLET, BC.OUTFLOW, BC.DMR_WALL, BC.DMR_TOP]
    bc_prim = np.zeros((6, 5))
    bc_prim[0, :4] = bc_prim[2, :4] = bc_prim[3, :4] = post
    s = Solver(2, n, h, InterMethod.TCNS5, flux, Recon.CHAR, bc + [BC.OUTFLOW] * 2, bc_prim, origin=x0)
    grid = oracle_grid(2, n, h, InterMethod.TCNS5, flux, Recon.CHAR, bc, bc_prim, x0=x0)
    _stepped_parity(s, grid, cons, 0.013, f"double Mach {flux.name}")


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_rk3_parity_gravity(dim):
    """The gravity source term on the last sweep, for any acceleration vector."""
    n = {1: (70,), 2: (36, 40), 3: (14, 12, 16)}[dim]
    gvec = [0.3, 1.0, -0.6][:dim]
    cons = random_state(dim, n, seed=17, jumps=True)
    bc = [BC.PERIODIC] * 2 + [BC.DIRICHLET] * 2 + [BC.SYMMETRYZ, BC.OUTFLOW]
    bc_prim = np.zeros((6, 5))
    bc_prim[2, :dim + 2] = [1.1] + [0.0] * dim + [1.3]
    bc_prim[3, :dim + 2] = [0.9] + [0.1] * dim + [0.8]
    s = Solver(dim, n, 0.04, InterMethod.WCNSZ5, FluxMethod.HLLC, Recon.CHAR, bc, bc_prim, gravity=gvec)
    grid = oracle_grid(dim, n, 0.04, InterMethod.WCNSZ5, FluxMethod.HLLC, Recon.CHAR, bc, bc_prim,
                       gravity=gvec + [0.0] * (3 - dim))
    _say(f"gravity dim={dim}")
    s.set_state(cons)
    _check(s.rhs(), O.rhs(grid, cons)[0], f"gravity rhs dim={dim}")
    _stepped_parity(s, grid, cons, 0.0, f"gravity rk3 dim={dim}")


def test_cfl_dt_parity():
    for dim in (1, 2, 3):
        n = SIZES[dim]
        cons = random_state(dim, n, seed=5)
        s = Solver(dim, n, [0.07, 0.05, 0.03][:dim], InterMethod.WCNS5, FluxMethod.ROE, Recon.CHAR)
        s.set_state(cons)
        g = s.cfl_dt(0.5)
        s.close()
        ref = O.cfl_dt(O.make_grid(dim, n, [0.07, 0.05, 0.03][:dim], O.S_JS, 1, 1, [7] * 6), cons, 0.5)
        assert g == ref, (dim, g, ref)


def test_vortex_256_rhs_and_steps_parity():
    cons, h = cases.vortex(256)
    bc = [BC.PERIODIC] * 4
    s = Solver(2, (256, 256), h, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, bc)
    s.set_state(cons)
    grid = oracle_grid(2, (256, 256), h, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, bc)
    ref, _ = O.rhs(grid, cons)
    _check(s.rhs(), ref, "vortex 256^2 rhs")
    s.rk3_steps_cfl(0.5, 2)
    ref2, _ = O.rk3(grid, cons, cfl=0.5, nsteps=2)
    _check(s.get_state(), ref2, "vortex 256^2 two CFL steps")
    s.close()


def test_full_size_conservation_2048():
    """Periodic vortex at 2048^2: totals of rho, rho u, rho v, rho E are
    conserved to round-off over 5 CFL steps (size-independent invariant)."""
    cons, h = cases.vortex(2048)
    s = Solver(2, (2048, 2048), h, InterMethod.WCNSZ5, FluxMethod.ROE, Recon.CHAR, [BC.PERIODIC] * 4)
    s.set_state(cons)
    t0 = s.totals()
    s.rk3_steps_cfl(0.5, 5)
    t1 = s.totals()
    s.close()
    scale = np.abs(cons).sum(0)
    _say("2048^2 totals drift", (t1 - t0) / scale)
    assert np.all(np.abs(t1 - t0) / scale < 1e-13)


def test_full_size_3d_freestream_and_conservation():
    n = 256
    cons, h = cases.periodic3d(n)
    s = Solver(3, (n, n, n), h, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6)
    s.set_state(cons)
    t0 = s.totals()
    s.rk3_steps_cfl(0.5, 2)
    t1 = s.totals()
    scale = np.abs(cons).sum(0)
    _say("256^3 totals drift", (t1 - t0) / scale)
    assert np.all(np.abs(t1 - t0) / scale < 1e-13)
    # uniform state: rhs identically zero
    u = np.tile(cons[:1], (n ** 3, 1))
    s.set_state(u)
    r = s.rhs()
    s.close()
    assert np.all(r == 0.0)


def _manual_rk3(grid, u0, dt):
    """One RK3 step written out (the oracle's stage expressions), returning the
    new state and the last stage's rhs (the reference logs its rho Linf)."""
    r = O.rhs(grid, u0)[0]
    u1 = u0 - dt * r
    grid.t = grid.t + dt
    r = O.rhs(grid, u1)[0]
    u2 = 0.75 * u0 - 0.25 * dt * r + 0.25 * u1
    grid.t = grid.t - dt / 2
    r = O.rhs(grid, u2)[0]
    u3 = 1.0 / 3.0 * u0 - 2.0 / 3.0 * dt * r + 2.0 / 3.0 * u2
    grid.t = grid.t - dt / 2
    return u3, r


@pytest.mark.parametrize("dim", [1, 2])
def test_steps_to_output_is_steps_loop_cfl(dim):
    """cfd_steps_to_output: the CFL dt clamped at output times, the steps, the
    clock and the (t, dt, rhoLinf) log of stepsLoopCFL, all on the device,
    against the oracle driven through the same loop; a call stops after the
    step that reaches the output time (the queued steps after it do nothing),
    or after max_steps steps, and output_t carries over between calls."""
    n = {1: (120,), 2: (40, 36)}[dim]
    bc = [BC.OUTFLOW] * 6
    cons = random_state(dim, n, seed=41, jumps=True)
    s = Solver(dim, n, 0.02, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, bc)
    s.set_state(cons)
    grid = oracle_grid(dim, n, 0.02, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, bc)
    out_dt = 0.0137
    u, t, out_t = cons.copy(), 0.0, 0.0
    log_all, calls = [], []
    for max_steps in (5, 64, 64, 3, 64):
        log, out_t_dev, reached = s.steps_to_output(0.5, out_dt, out_t, max_steps)
        calls.append((len(log), reached))
        for rec in log:
            dt = O.cfl_dt(grid, u, 0.5)
            hit = out_t + dt > out_dt
            if hit:
                dt = out_dt - out_t
                out_t = 0.0
            else:
                out_t += dt
            grid.t = t
            u, r = _manual_rk3(grid, u, dt)
            t = ((t + dt) - dt / 2 - dt / 2) + dt
            assert rec[1] == dt and rec[0] == t, (rec, dt, t)
            assert rec[2] == np.abs(r[:, 0]).max()
        assert out_t_dev == out_t
        assert reached == (out_t == 0.0)
        assert reached or len(log) == max_steps
        log_all += list(log)
    _say(f"steps_to_output dim={dim}: calls (steps, reached) {calls}")
    assert any(c[1] for c in calls) and any(not c[1] for c in calls)
    _check(s.get_state(), u, f"steps_to_output dim={dim}")
    assert s.get_time() == t
    s.close()
