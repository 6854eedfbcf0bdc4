"""Closed-form pins of the oracle's interpolation weights and Riemann fluxes (CPU).

Each check is chosen so that a plausible slip in the restated formula fails it:

* WENO-Z (weno5_Z, zone_FD/src/include/interScheme.hpp:85-114): with
  beta_0 == beta_2 its tau vanishes and it is the linear 5th-order stencil
  (a wrong tau, e.g. |beta_1 - beta_0|, does not vanish there); the deviation
  from the linear stencil falls like h^6 at a first-order critical point
  (f' = 0), where WENO-JS's falls only like h^4 (Henrick et al. 2005's order
  loss; Borges et al. 2008's fix) -- a linear tau/beta instead of its square,
  or wrong ideal weights, break these orders;
* WENO-JS (weno5_JSchen, :17-43): h^5 deviation at a regular point, h^4 at a
  critical point; its epsilon = 1e-6 makes it linear on data of amplitude
  well below 1e-3 and essentially ENO (the stencil across a jump weighted out)
  well above -- so an epsilon orders of magnitude off fails;
* Roe (roeFlux1D2 / roeFlux2DSym, fluxSchemes.hpp:13-71, 420-553): where the
  Harten fix is inactive the flux equals the upwind decomposition
  F(U_L) + sum_{lambda_k < 0} lambda_k alpha_k r_k = F(U_R) - sum_{lambda_k > 0} ...
  with alpha = L (U_R - U_L) from the oracle's Roe-averaged eigenvectors (pinned
  separately: L R = I, L A R diagonal) and lambda_k from Roe averages computed
  here -- the flux's own closed-form coefficients are never used, so a wrong
  coefficient, sign or wave speed in them fails;
* HLLC (HLLCFlux2D2, fluxSchemes.hpp:202-327; the 3D form): an isolated
  contact (density jump only, stationary or moving) and an isolated shear
  wave (tangential velocity jump only) at subsonic normal speed are resolved
  exactly: the flux is the physical flux of the upwind side.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

G = 1.4


def _dev_orders(scheme, f, off, h0=0.1, levels=8, amp=1e3):
    devs = []
    for k in range(levels):
        h = h0 / 2 ** k
        q = np.array([amp * f((j - 2) * h + off * h) for j in range(5)])
        devs.append(abs(O.inter5(scheme, q) - O.inter5(O.S_LINEAR, q)) / amp)
    return [math.log2(devs[i] / devs[i + 1]) for i in range(len(devs) - 1) if devs[i + 1] > 0]


def test_weno_z_is_linear_when_beta0_equals_beta2():
    rng = np.random.default_rng(7)
    for _ in range(200):
        a, b, c = rng.normal(size=3)
        q = np.array([a, b, c, b, a])  # beta_0 == beta_2 (mirror data about q_2)
        lin = O.inter5(O.S_LINEAR, q)
        scale = np.abs(q).max()
        assert abs(O.inter5(O.S_Z, q) - lin) <= 8e-16 * scale * 16
        # WENO-JS weighs by the betas themselves: not linear on the same data
        # unless beta_1 happens to equal the others
    q = np.array([0.0, 1.0, 3.0, 1.0, 0.0])
    assert abs(O.inter5(O.S_JS, q) - O.inter5(O.S_LINEAR, q)) > 1e-3


def test_weno_deviation_orders_at_critical_and_regular_points():
    crit = lambda x: 3 * x * x + x ** 3 + 10 * x ** 5  # f'(0) = 0, f'' != 0
    regular = lambda x: math.sin(2 * x) + 3 * x
    js_c = _dev_orders(O.S_JS, crit, 0.3)[2:]
    z_c = _dev_orders(O.S_Z, crit, 0.3)[2:]
    js_r = _dev_orders(O.S_JS, regular, 0.3)[1:]
    print("deviation orders: JS critical", js_c, "Z critical", z_c, "JS regular", js_r)
    assert all(3.8 < o < 4.3 for o in js_c), js_c
    assert all(o > 5.8 for o in z_c), z_c
    assert all(4.8 < o < 5.2 for o in js_r), js_r


def test_weno_js_epsilon_scale():
    step = np.array([0.0, 0.0, 0.0, 1.0, 1.0])
    lin = O.inter5(O.S_LINEAR, step)
    u0 = 0.375 * step[0] - 1.25 * step[1] + 1.875 * step[2]  # the candidate left of the jump
    small = 1e-6  # beta ~ 1e-12 << 1e-6: the linear weights
    assert abs(O.inter5(O.S_JS, small * step) / small - lin) < 1e-4
    big = 1e2  # beta ~ 1e4 >> 1e-6: the stencils across the jump weighted out
    assert abs(O.inter5(O.S_JS, big * step) / big - u0) < 1e-9
    # the transition sits where beta ~ epsilon, i.e. amplitude ~ 1e-3
    mid = 1e-3
    v = O.inter5(O.S_JS, mid * step) / mid
    assert abs(v - lin) > 0.05 and abs(v - u0) > 0.05


def _cons(p, dim):
    r, vel, pp = p[0], p[1:1 + dim], p[-1]
    return np.array([r, *(r * vel), pp / (G - 1) + r * (vel @ vel) / 2])


def _phys_flux(p, dim, ax):
    r, vel, pp = p[0], p[1:1 + dim], p[-1]
    vn = vel[ax]
    H = pp / r * G / (G - 1) + (vel @ vel) / 2
    mom = r * vel * vn
    mom[ax] += pp
    return np.array([r * vn, *mom, r * H * vn])


def _roe_waves(pl, pr, dim, ax):
    """Roe-averaged normal velocity and sound speed, written out here."""
    sl, sr = math.sqrt(pl[0]), math.sqrt(pr[0])
    vel = (sl * pl[1:1 + dim] + sr * pr[1:1 + dim]) / (sl + sr)
    Hl = pl[-1] / pl[0] * G / (G - 1) + (pl[1:1 + dim] @ pl[1:1 + dim]) / 2
    Hr = pr[-1] / pr[0] * G / (G - 1) + (pr[1:1 + dim] @ pr[1:1 + dim]) / 2
    H = (sl * Hl + sr * Hr) / (sl + sr)
    c = math.sqrt((G - 1) * (H - (vel @ vel) / 2))
    return vel[ax], c


def _lambdas(vn, c, dim):
    # characteristic order of the oracle's eigensystems: 1D (Vn-c, Vn, Vn+c);
    # 2D (shear, entropy, Vn-c, Vn+c); 3D adds a second shear wave last
    if dim == 1:
        return np.array([vn - c, vn, vn + c])
    lam = [vn, vn, vn - c, vn + c]
    if dim == 3:
        lam.append(vn)
    return np.array(lam)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_roe_flux_is_the_upwind_decomposition_where_the_fix_is_inactive(dim):
    rng = np.random.default_rng(100 + dim)
    nv = dim + 2
    fix = 0.1 if dim == 1 else 0.05  # Harten constants of roeFlux1D2 / roeFlux2DSym
    checked = 0
    while checked < 60:
        pl = np.array([rng.uniform(0.3, 2.0), *rng.uniform(-1.5, 1.5, dim), rng.uniform(0.3, 2.0)])
        pr = np.array([rng.uniform(0.3, 2.0), *rng.uniform(-1.5, 1.5, dim), rng.uniform(0.3, 2.0)])
        for ax in range(dim):
            vn, c = _roe_waves(pl, pr, dim, ax)
            lam = _lambdas(vn, c, dim)
            if np.any(np.abs(lam) < fix * (abs(vn) + c) * 1.01):
                continue  # the entropy fix would act: not this check
            L, R = O.eigen(dim, pl, pr, ax)
            alpha = L @ (_cons(pr, dim) - _cons(pl, dim))
            up = _phys_flux(pl, dim, ax) + R @ np.where(lam < 0, lam * alpha, 0.0)
            dn = _phys_flux(pr, dim, ax) - R @ np.where(lam > 0, lam * alpha, 0.0)
            got = O.riemann(dim, O.F_ROE, pl, pr, ax)
            scale = np.abs(_phys_flux(pl, dim, ax)).max() + np.abs(_phys_flux(pr, dim, ax)).max()
            assert np.abs(got - up).max() <= 1e-13 * scale, (dim, ax, got, up)
            assert np.abs(got - dn).max() <= 1e-13 * scale
            checked += 1
    # the identity separates: the same sum with the sign of the wave strengths
    # flipped (a plausible slip) is far from the flux
    pl = np.array([1.0, *([0.3] * dim), 1.0])
    pr = np.array([0.5, *([0.1] * dim), 0.4])
    L, R = O.eigen(dim, pl, pr, 0)
    vn, c = _roe_waves(pl, pr, dim, 0)
    lam = _lambdas(vn, c, dim)
    alpha = L @ (_cons(pr, dim) - _cons(pl, dim))
    wrong = _phys_flux(pl, dim, 0) - R @ np.where(lam < 0, lam * alpha, 0.0)
    assert np.abs(O.riemann(dim, O.F_ROE, pl, pr, 0) - wrong).max() > 1e-3


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_resolves_isolated_contacts_and_shear_waves(dim):
    rng = np.random.default_rng(200 + dim)
    nv = dim + 2
    for trial in range(120):
        ax = trial % dim
        p = rng.uniform(0.2, 2.0)
        rl, rr = rng.uniform(0.2, 2.0, 2)
        vel = rng.uniform(-1.0, 1.0, dim)
        c = math.sqrt(G * p / max(rl, rr))
        kind = trial % 3
        vel[ax] = 0.0 if kind == 0 else rng.uniform(-0.6, 0.6) * c  # stationary / moving
        vl, vr = vel.copy(), vel.copy()
        if kind == 2:  # shear wave: tangential velocities jump, density equal
            for d in range(dim):
                if d != ax:
                    vr[d] = vel[d] + rng.uniform(-1, 1)
            rr = rl
        pl = np.array([rl, *vl, p])
        pr = np.array([rr, *vr, p])
        got = O.riemann(dim, O.F_HLLC, pl, pr, ax)
        up = pl if vel[ax] > 0 else pr
        want = _phys_flux(up, dim, ax)
        if vel[ax] == 0.0:
            want = (_phys_flux(pl, dim, ax) + _phys_flux(pr, dim, ax)) / 2
        scale = max(np.abs(_phys_flux(pl, dim, ax)).max(), np.abs(_phys_flux(pr, dim, ax)).max())
        assert np.abs(got - want).max() <= 2e-14 * scale, (dim, ax, kind, got, want)
    # a planted slip fails: the contact's density jump seen by a flux that
    # averages the two sides (a central flux) is not the upwind flux
    pl = np.array([1.0, *([0.3] + [0.0] * (dim - 1)), 1.0])
    pr = np.array([0.2, *([0.3] + [0.0] * (dim - 1)), 1.0])
    central = (_phys_flux(pl, dim, 0) + _phys_flux(pr, dim, 0)) / 2
    assert np.abs(O.riemann(dim, O.F_HLLC, pl, pr, 0) - central).max() > 1e-3


def test_entropy_fix_acts_in_the_reference_roe_runs():
    """The Roe runs reproduced bit for bit (CongFD1.0's Lax series and
    Rayleigh-Taylor files) exercise the Harten fix: on the RTI initial state
    the entropy and shear waves of every x interface (u = 0) and of most y
    interfaces (|v| << c) fall below 0.05 (|Vn| + c)."""
    from tests.test_reference_outputs_congfd1 import _rti_initial
    cons, _ = _rti_initial(5.0 / 3.0)
    g = 5.0 / 3.0
    c3 = cons.reshape(256, 64, 4)
    r = c3[..., 0]
    v = c3[..., 2] / r
    p = (g - 1) * (c3[..., 3] - 0.5 * r * v * v)
    c = np.sqrt(g * p / r)
    frac_y = np.mean(np.abs(v) < 0.05 * (np.abs(v) + c))
    assert frac_y > 0.9


def _betas(q):
    b0 = (q[0] - 2 * q[1] + q[2]) ** 2 + 0.25 * (q[0] - 4 * q[1] + 3 * q[2]) ** 2
    b1 = (q[1] - 2 * q[2] + q[3]) ** 2 + 0.25 * (q[1] - q[3]) ** 2
    b2 = (q[2] - 2 * q[3] + q[4]) ** 2 + 0.25 * (3 * q[2] - 4 * q[3] + q[4]) ** 2
    return b0, b1, b2


@pytest.mark.parametrize("scheme", [O.S_TENO5, O.S_TENO5_CT4, O.S_TENO5_CT7, O.S_TENO5_CT10, O.S_TENOA])
def test_teno_keeps_every_stencil_when_tau_is_below_the_smallest_beta(scheme):
    """The GPU kernels skip the gamma-ratio cut-off (three divisions and the
    smoothness ratios of TENO-A) where tau = |beta_2 - beta_0| <= min beta_k +
    1e-40, returning flag 0: 1 <= gamma_k <= 64 there and ct * sum(gamma) <=
    192 ct < 1 for every constant of the family.  The oracle's full
    evaluation agrees on smooth, scaled, tied (tau == min beta) and random
    windows: the value is the 5th-order linear stencil's, bit for bit."""
    rng = np.random.default_rng(int(scheme))
    hits = 0
    for trial in range(20000):
        kind = trial % 4
        if kind == 0:  # smooth samples of random functions
            h = 10 ** rng.uniform(-4, -0.5)
            a, b, c, d = rng.normal(size=4)
            x0 = rng.uniform(-1, 1)
            q = np.array([a + b * math.sin(c * (x0 + k * h)) + d * (x0 + k * h) ** 3 for k in range(5)])
        elif kind == 1:  # random windows scaled over many decades
            q = rng.normal(size=5) * 10 ** rng.uniform(-30, 30)
        elif kind == 2:  # near-constant data (beta ~ 0)
            q = 1.0 + rng.normal(size=5) * 10 ** rng.uniform(-300, -8)
        else:  # mirrored data: beta_0 == beta_2, tau == 0
            a, b, c = rng.normal(size=3)
            q = np.array([a, b, c, b, a])
        b0, b1, b2 = _betas(q)
        tau = abs(b2 - b0)
        if tau <= min(b0, b1, b2) + 1e-40:
            hits += 1
            assert O.inter5(scheme, q) == O.inter5(O.S_LINEAR, q), q
    assert hits > 5000
