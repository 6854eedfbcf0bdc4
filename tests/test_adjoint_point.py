"""CPU checks of the hand-written adjoint point kernels (adjoint_point.hpp):
every vector-Jacobian product against central finite differences of its own
forward function, and the adjoint's forward model against the oracle's
interface flux (so the adjoint differentiates the same map the solver runs)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "adjcheck", "libadjcheck.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "adjcheck")], check=True)
    L = ctypes.CDLL(LIB)
    pd = ctypes.POINTER(ctypes.c_double)
    L.adj_interface_fwd.argtypes = [ctypes.c_int, pd, pd]
    L.adj_interface_vjp.argtypes = [ctypes.c_int, pd, pd, pd]
    L.adj_interface_fwd_s.argtypes = [ctypes.c_int, pd, pd, ctypes.c_int, ctypes.c_double]
    L.adj_interface_vjp_s.argtypes = [ctypes.c_int, pd, pd, pd, ctypes.c_int, ctypes.c_double]
    L.adj_hllc_fwd.argtypes = [ctypes.c_int, pd, pd, pd]
    L.adj_hllc_vjp.argtypes = [ctypes.c_int, pd, pd, pd, pd, pd]
    L.adj_js_fwd.argtypes = [pd]
    L.adj_js_fwd.restype = ctypes.c_double
    L.adj_js_vjp.argtypes = [pd, ctypes.c_double, pd]
    L.adj_node_flux_vjp.argtypes = [ctypes.c_int, pd, pd, pd]
    L.adj_node_flux.argtypes = [ctypes.c_int, pd, pd]
    L.adj_c2p_vjp.argtypes = [pd, pd, pd]
    return L


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fd_grad(f, x, lam, eps=1e-6):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = eps * max(1.0, abs(x.flat[i]))
        g.flat[i] = (lam @ (f(x + e) - f(x - e))) / (2 * e.flat[i])
    return g


def rand_prim(rng, n):
    w = np.empty((n, 4))
    w[:, 0] = rng.uniform(0.5, 1.5, n)
    w[:, 1] = rng.uniform(-0.8, 0.8, n)
    w[:, 2] = rng.uniform(-0.8, 0.8, n)
    w[:, 3] = rng.uniform(0.5, 1.5, n)
    return w


def test_js_vjp(lib):
    rng = np.random.default_rng(1)
    for _ in range(20):
        q = rng.normal(size=5)
        g = np.zeros(5)
        lib.adj_js_vjp(P(q), 1.0, P(g))
        out = np.zeros(5)
        lib.adj_js_vjp(P(q), 1.0, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        f = lambda x: np.array([lib.adj_js_fwd(P(x))])
        np.testing.assert_allclose(g, fd_grad(f, q, np.ones(1)), rtol=1e-6, atol=1e-8)
        # forward equals the oracle's WENO-JS
        assert abs(lib.adj_js_fwd(P(q)) - O.inter5(O.S_JS, q)) < 1e-13 * (1 + abs(O.inter5(O.S_JS, q)))


@pytest.mark.parametrize("ax", [0, 1])
def test_hllc_vjp_all_branches(lib, ax):
    rng = np.random.default_rng(2 + ax)
    seen = set()
    for trial in range(200):
        L, R = rand_prim(rng, 2)
        kind = trial % 4
        if kind == 1:   # supersonic right-moving
            L[1 + ax], R[1 + ax] = 3.0 + rng.uniform(), 3.0 + rng.uniform()
        elif kind == 2:  # supersonic left-moving
            L[1 + ax], R[1 + ax] = -3.0 - rng.uniform(), -3.0 - rng.uniform()
        elif kind == 3:  # strong compression (shock factors active)
            L[1 + ax], R[1 + ax] = 1.5, -1.5
        lF = rng.normal(size=4)
        lL, lR = np.zeros(4), np.zeros(4)
        lib.adj_hllc_vjp(ax, P(L), P(R), P(lF), lL.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         lR.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

        def f(x):
            F = np.zeros(4)
            lib.adj_hllc_fwd(ax, P(x[:4]), P(x[4:]), F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            return F

        x = np.concatenate([L, R])
        g = fd_grad(f, x, lF, eps=1e-7)
        np.testing.assert_allclose(np.concatenate([lL, lR]), g, rtol=2e-5, atol=2e-7)
        # forward equals the oracle's HLLCFlux2D2
        F = f(x)
        np.testing.assert_allclose(F, O.riemann(2, O.F_HLLC, L, R, ax), rtol=1e-13, atol=1e-13)
        Fo = O.riemann(2, O.F_HLLC, L, R, ax)
        seen.add(kind)
    assert seen == {0, 1, 2, 3}


@pytest.mark.parametrize("ax", [0, 1])
def test_node_flux_and_c2p_vjp(lib, ax):
    rng = np.random.default_rng(5)
    for _ in range(10):
        p = rand_prim(rng, 1)[0]
        lF = rng.normal(size=4)
        g = np.zeros(4)
        lib.adj_node_flux_vjp(ax, P(p), P(lF), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

        def f(x):
            F = np.zeros(4)
            lib.adj_node_flux(ax, P(x), F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            return F

        np.testing.assert_allclose(g, fd_grad(f, p, lF), rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose(f(p), O.node_flux(2, p, ax), rtol=1e-14)
        c = np.array([p[0], p[0] * p[1], p[0] * p[2], p[3] / 0.4 + 0.5 * p[0] * (p[1] ** 2 + p[2] ** 2)])
        lp = rng.normal(size=4)
        gc = np.zeros(4)
        lib.adj_c2p_vjp(P(c), P(lp), gc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        np.testing.assert_allclose(gc, fd_grad(lambda x: O.cons_to_prim(2, x), c, lp), rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("ax", [0, 1])
@pytest.mark.parametrize("kind", ["smooth", "jump"])
def test_interface_vjp(lib, ax, kind):
    rng = np.random.default_rng(7 + ax)
    for trial in range(15):
        w = rand_prim(rng, 6)
        if kind == "smooth":
            base = rand_prim(rng, 1)[0]
            w = base + 0.05 * (w - w.mean(0))
        else:
            w[3:, 0] *= 0.3
            w[3:, 3] *= 0.2
        lF = rng.normal(size=4)
        lw = np.zeros(24)
        lib.adj_interface_vjp(ax, P(w), P(lF), lw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

        def f(x):
            F = np.zeros(4)
            lib.adj_interface_fwd(ax, P(x), F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            return F

        x = w.reshape(-1)
        np.testing.assert_allclose(lw, fd_grad(f, x, lF, eps=1e-7), rtol=5e-5, atol=1e-6)
        # the forward model is the solver's: oracle recon (JS, char) + HLLC
        Lr, Rr = O.recon(2, O.S_JS, 1, w, ax)
        np.testing.assert_allclose(f(x), O.riemann(2, O.F_HLLC, Lr, Rr, ax), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("ax", [0, 1])
@pytest.mark.parametrize("kind", ["smooth", "jump"])
def test_interface_vjp_teno5(lib, ax, kind):
    """The TENO5 interface pass: its interpolation is the flag's fixed
    stencil (linear), so the gradient is the stencil's; against central
    differences of the same forward (flags do not move under the 1e-7
    step on these states), and the forward against the oracle's TENO5
    reconstruction + HLLC."""
    rng = np.random.default_rng(17 + ax)
    cuts = 0
    for trial in range(15):
        w = rand_prim(rng, 6)
        if kind == "smooth":
            base = rand_prim(rng, 1)[0]
            w = base + 0.05 * (w - w.mean(0))
        else:
            w[3:, 0] *= 0.3
            w[3:, 3] *= 0.2
        lF = rng.normal(size=4)
        lw = np.zeros(24)
        lib.adj_interface_vjp_s(ax, P(w), P(lF), lw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, 1e-5)

        def f(x):
            F = np.zeros(4)
            lib.adj_interface_fwd_s(ax, P(x), F.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, 1e-5)
            return F

        x = w.reshape(-1)
        np.testing.assert_allclose(lw, fd_grad(f, x, lF, eps=1e-7), rtol=5e-5, atol=1e-6)
        Lr, Rr = O.recon(2, O.S_TENO5, 1, w, ax)
        np.testing.assert_allclose(f(x), O.riemann(2, O.F_HLLC, Lr, Rr, ax), rtol=1e-10, atol=1e-12)
        Lj, _ = O.recon(2, O.S_JS, 1, w, ax)
        cuts += not np.allclose(Lr, Lj, rtol=1e-3)
    if kind == "jump":
        assert cuts > 5  # TENO cuts stencils here: a JS pass would not match
