"""Initial conditions of the workloads (host side, numpy).

All functions return the conservative state in the reference layout:
``(nz, ny, nx, nv)`` flattened to ``[cell][nv]`` with x fastest, plus the
cell spacing.  Variables: (rho, rho u[, rho v[, rho w]], rho E), gamma = 1.4.
"""
import math

import numpy as np

GAMMA = 1.4


def prim_to_cons(prim):
    """(rho, u.., p) -> (rho, rho u.., p/(gamma-1) + rho |u|^2 / 2) along the last axis."""
    prim = np.asarray(prim, dtype=np.float64)
    d = prim.shape[-1] - 2
    r = prim[..., 0]
    vel = prim[..., 1:1 + d]
    p = prim[..., -1]
    out = np.empty_like(prim)
    out[..., 0] = r
    out[..., 1:1 + d] = r[..., None] * vel
    out[..., -1] = p / (GAMMA - 1) + 0.5 * r * np.sum(vel * vel, axis=-1)
    return out


def cons_to_prim(cons):
    cons = np.asarray(cons, dtype=np.float64)
    d = cons.shape[-1] - 2
    r = cons[..., 0]
    vel = cons[..., 1:1 + d] / r[..., None]
    out = np.empty_like(cons)
    out[..., 0] = r
    out[..., 1:1 + d] = vel
    out[..., -1] = (GAMMA - 1) * (cons[..., -1] - 0.5 * r * np.sum(vel * vel, axis=-1))
    return out


def sod(n=400):
    """1D Sod tube on [0,1], split at 0.5: (1,0,1) | (0.125,0,0.1).

    Same states as the reference preset (zone_FD/src/src/initializer.cpp:79-103,
    energy written as 1/(gamma-1)*p)."""
    h = 1.0 / n
    x = (np.arange(n) + 0.5) * h
    c = np.empty((n, 3))
    left = x < 0.5
    c[:, 0] = np.where(left, 1.0, 0.125)
    c[:, 1] = 0.0
    c[:, 2] = np.where(left, 1.0 / (GAMMA - 1) * 1, 1.0 / (GAMMA - 1) * 0.1)
    return c, h


def vortex_prim(x, y, t=0.0, L=10.0, eps=5.0, xc=5.0, yc=5.0, u0=1.0, v0=1.0):
    """Isentropic vortex (Shu) advected with (u0, v0) on a periodic [0,L]^2."""
    xs = np.mod(x - u0 * t - xc + L / 2, L) - L / 2
    ys = np.mod(y - v0 * t - yc + L / 2, L) - L / 2
    r2 = xs * xs + ys * ys
    f = np.exp(0.5 * (1.0 - r2))
    du = -eps / (2 * math.pi) * f * ys
    dv = eps / (2 * math.pi) * f * xs
    T = 1.0 - (GAMMA - 1) * eps * eps / (8 * GAMMA * math.pi ** 2) * f * f
    rho = T ** (1.0 / (GAMMA - 1))
    p = rho ** GAMMA
    return np.stack([rho, u0 + du, v0 + dv, p], -1)


def vortex(n, t=0.0, L=10.0):
    """2D isentropic vortex on a periodic n x n grid of [0,L]^2 (cell centres)."""
    h = L / n
    xc = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(xc, xc, indexing="xy")  # [y][x]
    prim = vortex_prim(X, Y, t, L)
    return prim_to_cons(prim).reshape(-1, 4), h


def _random_phase_field(n, kmax, rng):
    """Real field on an n^3 periodic grid with unit-amplitude random-phase
    Fourier modes 1 <= |k| <= kmax, scaled to max |f| = 1."""
    k = np.fft.fftfreq(n, d=1.0 / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, kr, indexing="ij")
    kk = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    mask = (kk >= 1) & (kk <= kmax)
    phase = rng.uniform(0, 2 * np.pi, size=kk.shape)
    spec = np.where(mask, np.exp(1j * phase), 0.0)
    f = np.fft.irfftn(spec, s=(n, n, n), axes=(0, 1, 2))
    return f / np.abs(f).max()


def periodic3d(n, seed=20261203, mach=0.3, amp=0.05, kmax=8):
    """3D periodic flow on [0,2pi]^3: rho = 1 + amp*f_rho, (u,v,w) = amp*f_k,
    p = 1/(gamma M^2), with seeded random-phase modes |k| <= kmax."""
    rng = np.random.default_rng(seed)
    h = 2 * math.pi / n
    prim = np.empty((n, n, n, 5))
    prim[..., 0] = 1.0 + amp * _random_phase_field(n, kmax, rng)
    for v in (1, 2, 3):
        prim[..., v] = amp * _random_phase_field(n, kmax, rng)
    prim[..., 4] = 1.0 / (GAMMA * mach * mach)
    return prim_to_cons(prim).reshape(-1, 5), h


def _random_phase_field_2d(n, kmax, rng):
    """Real field on an n^2 periodic grid ([y][x]) with unit-amplitude
    random-phase Fourier modes 1 <= |k| <= kmax, scaled to max |f| = 1."""
    k = np.fft.fftfreq(n, d=1.0 / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n)
    KY, KX = np.meshgrid(k, kr, indexing="ij")
    kk = np.sqrt(KX * KX + KY * KY)
    mask = (kk >= 1) & (kk <= kmax)
    phase = rng.uniform(0, 2 * np.pi, size=kk.shape)
    spec = np.where(mask, np.exp(1j * phase), 0.0)
    f = np.fft.irfftn(spec, s=(n, n), axes=(0, 1))
    return f / np.abs(f).max()


# Rankine-Hugoniot jumps of a normal shock at upstream Mach 1.5, gamma 1.4
_M_SHOCK = 1.5
_RHO_JUMP = (GAMMA + 1) * _M_SHOCK ** 2 / ((GAMMA - 1) * _M_SHOCK ** 2 + 2)
_P_JUMP = 1 + 2 * GAMMA / (GAMMA + 1) * (_M_SHOCK ** 2 - 1)


def adjoint_shock(n, seed=20261204, amp=1e-2, kmax=8):
    """Adjoint workload on the periodic unit square, n x n cells.

    rho0 = 1 + amp * f (seeded random-phase modes |k| <= kmax) multiplies a
    density profile that holds a stationary normal shock at x = 1/2: upstream
    (rho, u, p) = (1, 1.5, 1/gamma) (Mach 1.5), downstream the
    Rankine-Hugoniot state, which relaxes smoothly back to the upstream state
    around x = 0.9 so that the periodic seam is not a second shock; v = 0.
    Returns a dict with rho0 [ny][nx], the profile factors (rho_factor, u,
    p, all [ny][nx]), h, and the objective window (y slice, x slice) of
    n/8 x n/8 cells just downstream of the shock.
    """
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    phi = np.where(xc > 0.5, 1.0, 0.0) * 0.5 * (1.0 - np.tanh((xc - 0.9) / 0.02))
    u_up, p_up = _M_SHOCK, 1.0 / GAMMA
    u_dn = u_up / _RHO_JUMP
    rho_factor = 1.0 + (_RHO_JUMP - 1.0) * phi
    u = u_up + (u_dn - u_up) * phi
    p = p_up * (1.0 + (_P_JUMP - 1.0) * phi)
    rho0 = 1.0 + amp * _random_phase_field_2d(n, kmax, rng)
    w = max(n // 8, 1)
    i0 = n // 2 + max(n // 128, 1)
    j0 = n // 2 - w // 2
    ones = np.ones((n, 1))
    return {"rho0": rho0, "rho_factor": ones * rho_factor, "u": ones * u, "v": np.zeros((n, n)), "p": ones * p,
            "h": h, "window": (slice(j0, j0 + w), slice(i0, i0 + w))}


def adjoint_smooth(n, seed=20261205, amp=1e-2, kmax=8, u0=0.5, v0=0.3):
    """Smooth companion of adjoint_shock on the periodic unit square: a
    uniform subsonic stream (u0, v0) at p = 1/gamma, rho = 1 (Mach 0.58)
    carrying seeded random-phase perturbations of amplitude amp (|k| <= kmax)
    in all four primitive variables -- acoustic and entropy waves with no
    discontinuity, so the discrete map has no branch switch a small
    perturbation could cross (no shock, no TENO flags with WENO-JS, HLLC
    always in its subsonic star region with S* > 0).  Same dict layout and
    objective window as adjoint_shock; rho0 holds the density perturbation
    (rho_factor = 1)."""
    rng = np.random.default_rng(seed)
    rho0 = 1.0 + amp * _random_phase_field_2d(n, kmax, rng)
    u = u0 + amp * _random_phase_field_2d(n, kmax, rng)
    v = v0 + amp * _random_phase_field_2d(n, kmax, rng)
    p = 1.0 / GAMMA * (1.0 + amp * _random_phase_field_2d(n, kmax, rng))
    w = max(n // 8, 1)
    i0 = n // 2 + max(n // 128, 1)
    j0 = n // 2 - w // 2
    return {"rho0": rho0, "rho_factor": np.ones((n, n)), "u": u, "v": v, "p": p, "h": 1.0 / n,
            "window": (slice(j0, j0 + w), slice(i0, i0 + w))}


def adjoint_state(case, rho0=None):
    """Conservative state [cell][4] of an adjoint_shock case for the density field rho0."""
    r = (case["rho0"] if rho0 is None else rho0) * case["rho_factor"]
    prim = np.stack([r, case["u"], case["v"], case["p"]], -1)
    return prim_to_cons(prim).reshape(-1, 4)


def cfl_dt_2d(cons, n, h, cfl=0.5):
    """dt = CFL / sum_d max (c + |u_d|) / h (BlockSolver::getTimeIntervalExplicit)."""
    pr = cons_to_prim(np.asarray(cons).reshape(-1, 4))
    c = np.sqrt(GAMMA * pr[:, 3] / pr[:, 0])
    return cfl / ((c + np.abs(pr[:, 1])).max() / h + (c + np.abs(pr[:, 2])).max() / h)


def riemann2d_quadrants(nx, ny, seed):
    """2D four-quadrant Riemann problem on [0,1]^2 with quadrant states drawn
    from a seed: rho, p in [0.1, 1.5], |u|, |v| <= 1."""
    rng = np.random.default_rng(seed)
    q = np.empty((4, 4))
    q[:, 0] = rng.uniform(0.1, 1.5, 4)
    q[:, 1] = rng.uniform(-1, 1, 4)
    q[:, 2] = rng.uniform(-1, 1, 4)
    q[:, 3] = rng.uniform(0.1, 1.5, 4)
    h = 1.0 / max(nx, ny)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="xy")
    quad = np.where(X >= 0.5, np.where(Y >= 0.5, 0, 3), np.where(Y >= 0.5, 1, 2))
    prim = q[quad]
    return prim_to_cons(prim).reshape(-1, 4), h


def sod_exact(x, t, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1), x0=0.5):
    """Exact Riemann solution (Toro, ch. 4) of the Sod problem at time t."""
    g = GAMMA
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)

    def f(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-8, 10.0
    for _ in range(200):
        pm = 0.5 * (lo + hi)
        if f(pm, rl, pl, cl) + f(pm, rr, pr, cr) + (ur - ul) > 0:
            hi = pm
        else:
            lo = pm
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (f(ps, rr, pr, cr) - f(ps, rl, pl, cl))
    out = np.empty((len(x), 3))
    for i, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:
            if ps > pl:  # left shock
                sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s < sl:
                    out[i] = (rl, ul, pl)
                else:
                    out[i] = (rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1)), us, ps)
            else:  # left rarefaction
                shl = ul - cl
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - csl
                if s < shl:
                    out[i] = (rl, ul, pl)
                elif s > stl:
                    out[i] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[i] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:  # right shock
                sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s > sr:
                    out[i] = (rr, ur, pr)
                else:
                    out[i] = (rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)), us, ps)
            else:
                shr = ur + cr
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                st = us + csr
                if s > shr:
                    out[i] = (rr, ur, pr)
                elif s < st:
                    out[i] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[i] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out


def reference_1d(ncase, x0, x1, npoints, gamma=1.4):
    """Initial conservative state of the reference's 1D Euler presets
    (Initializer::solInit, zone_FD/src/src/initializer.cpp:79-242), on the
    cell centres x0 + (i + 1/2) h, h = (x1 - x0) / (npoints - 1), with the
    reference's own expressions (libm sin, the same operation order), so the
    state is the reference's bit for bit.  Cases: 0 Sod, 1 Shu-Osher, 2 Lax,
    3 Sedov, 4 Woodward-Colella, 5 double rarefaction.  Returns (cons, h)."""
    n = npoints - 1
    h = (x1 - x0) / (npoints - 1)
    G = gamma
    xc = [x0 + (i + 0.5) * h for i in range(n)]
    out = []
    for i, x in enumerate(xc):
        if ncase == 0:
            out.append([1, 0, 1.0 / (G - 1) * 1] if x < 0 else [0.125, 0, 1.0 / (G - 1) * 0.1])
        elif ncase == 1:
            if x <= 0.5:
                out.append([3.857143, 3.857143 * 2.629369,
                            1.0 / (G - 1.0) * (10.0 + 1.0 / 3.0) + 3.857143 * 2.629369 * 2.629369 / 2])
            else:
                out.append([1.0 + 0.2 * math.sin(5.0 * x), 0.0, 1.0 / (G - 1.0) * 1.0])
        elif ncase == 2:
            r, u, p = (0.445, 0.698, 3.528) if x <= 0 else (0.5, 0.0, 0.571)
            out.append([r, r * u, 1.0 / (G - 1) * p + r * u * u / 2])
        elif ncase == 3:
            E = 3200000.0 / (xc[i] - (x0 + (i - 1 + 0.5) * h)) if abs(x) < 1e-10 else 1e-12
            out.append([1, 0, 1 * E])
        elif ncase == 4:
            p = 1000 if x < 0.1 else (100 if x >= 0.9 else 0.01)
            out.append([1, 0, 1.0 / (G - 1) * p])
        elif ncase == 5:
            u = -2 if x < 0 else 2
            out.append([1, u, 1.0 / (G - 1) * 0.4 + u * u / 2])
        else:
            raise ValueError(f"1D case {ncase} is not defined")
    return np.array(out, dtype=np.float64), h
