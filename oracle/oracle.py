"""ctypes view of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: the test suite, the smoke check and the
benchmark's CPU-baseline leg use this module as the checker.  The product
package never imports it.

The oracle restates the reference solver's right-hand side and SSP-RK3 step
(wanghongfei23/CFD_inWSL, ``zone_FD/src``) in the reference's evaluation
order; see ``oracle.hpp`` for the file:line each function follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# oracle::Scheme
S_JS, S_Z, S_TENO5, S_TENO5_CT4, S_TENO5_CT7, S_TENOS, S_LINEAR = range(7)
S_TENOS_CT4, S_TENOS_CT7, S_TENO5_CT10, S_TENOA = range(7, 11)
S_TENOSA_F002, S_TENOSA_F203, S_TENOSA_F202 = range(11, 14)
S_WHF_N, S_WHF_A, S_WHF_AS, S_WHF_S, S_WHF_LAD = range(14, 19)
S_AI_S, S_AI_A, S_AI_AS, S_AI_LADS, S_AI_AZS, S_AI_F002, S_WHF_COMPARE = range(19, 26)
S_CONGAA, S_MUSCL5 = range(26, 28)
# CongFD1.0's whf_TCNS_AS_myF203_NoS and whf_TCNS_A
S_C1_F203, S_C1_TENOA = range(28, 30)
# oracle::Flux
F_HLLC, F_ROE = 0, 1


class OracleGrid(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("h", ctypes.c_double * 3),
        ("scheme", ctypes.c_int),
        ("flux", ctypes.c_int),
        ("charwise", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("bc_prim", (ctypes.c_double * 5) * 6),
        ("x0", ctypes.c_double),
        ("t", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("variant", ctypes.c_int),
        ("gamma_num", ctypes.c_double),
        ("gamma_den", ctypes.c_double),
    ]


# oracle Grid.variant
V_ZONE_FD, V_CONGFD1 = 0, 1
# the GAMMA macro's text: (num, den) for "num/den", (num, 0) for one literal
GAMMA_14 = (1.4, 0.0)
GAMMA_5_3 = (5.0, 3.0)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        d = ctypes.c_double
        pd = ctypes.POINTER(ctypes.c_double)
        i = ctypes.c_int
        pg = ctypes.POINTER(OracleGrid)
        _lib.oracle_rhs.argtypes = [pg, pd, pd, pd, i]
        _lib.oracle_rk3.argtypes = [pg, pd, d, d, i, i, pd]
        _lib.oracle_cfl_dt.argtypes = [pg, pd, d, i]
        _lib.oracle_cfl_dt.restype = d
        _lib.oracle_inter5.argtypes = [i, pd]
        _lib.oracle_inter5.restype = d
        _lib.oracle_cons_to_prim.argtypes = [i, pd, pd]
        _lib.oracle_eigen.argtypes = [i, pd, pd, i, pd, pd]
        _lib.oracle_prim_char_roundtrip.argtypes = [i, pd, pd, i, pd, pd, pd]
        _lib.oracle_riemann.argtypes = [i, i, pd, pd, i, pd]
        _lib.oracle_node_flux.argtypes = [i, pd, i, pd]
        _lib.oracle_recon.argtypes = [i, i, i, pd, i, pd, pd]
        _lib.oracle_accuracy_rhs.argtypes = [i, i, pd, pd, pd, d, pd]
        _lib.oracle_max_threads.restype = i
        _lib.oracle_set_gas.argtypes = [d, d]
        _lib.oracle_dual_rhs.argtypes = [pg, pd, pd, pd, pd, i]
        _lib.oracle_dual_rk3.argtypes = [pg, pd, pd, d, i, i]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_grid(dim, n, h, scheme, flux, charwise, bc, bc_prim=None, x0=0.0, t=0.0, gravity=(0.0, 0.0, 0.0),
              variant=V_ZONE_FD, gamma=GAMMA_14):
    g = OracleGrid()
    g.variant = int(variant)
    g.gamma_num, g.gamma_den = float(gamma[0]), float(gamma[1])
    g.x0 = float(x0)
    g.t = float(t)
    for i in range(3):
        g.gravity[i] = float(gravity[i]) if i < len(gravity) else 0.0
    g.dim = dim
    nn = list(n) + [1] * (3 - len(n))
    hh = list(h) if hasattr(h, "__len__") else [h] * 3
    hh = hh + [hh[-1]] * (3 - len(hh))
    for i in range(3):
        g.n[i] = int(nn[i])
        g.h[i] = float(hh[i])
    g.scheme = scheme
    g.flux = flux
    g.charwise = int(charwise)
    for i in range(6):
        g.bc[i] = int(bc[i]) if i < len(bc) else 7
        for v in range(5):
            g.bc_prim[i][v] = float(bc_prim[i][v]) if bc_prim is not None else 0.0
    return g


def nthreads_default():
    return lib().oracle_max_threads()


def rhs(grid, cons, nthreads=None):
    cons = _f64(cons)
    prim = np.empty_like(cons)
    out = np.empty_like(cons)
    rc = lib().oracle_rhs(ctypes.byref(grid), _p(cons), _p(prim), _p(out), nthreads or nthreads_default())
    assert rc == 0
    return out, prim


def rk3(grid, cons, dt=0.0, cfl=0.0, nsteps=1, nthreads=None):
    c = _f64(cons).copy()
    t = ctypes.c_double(0)
    rc = lib().oracle_rk3(ctypes.byref(grid), _p(c), float(dt), float(cfl), int(nsteps),
                          nthreads or nthreads_default(), ctypes.byref(t))
    assert rc == 0
    return c, t.value


def dual_rhs(grid, cons, dcons, nthreads=None):
    """rhs and its directional derivative along dcons (forward-mode duals)."""
    cons, dcons = _f64(cons), _f64(dcons)
    assert cons.shape == dcons.shape
    out, dout = np.empty_like(cons), np.empty_like(cons)
    rc = lib().oracle_dual_rhs(ctypes.byref(grid), _p(cons), _p(dcons), _p(out), _p(dout),
                               nthreads or nthreads_default())
    assert rc == 0
    return out, dout


def dual_rk3(grid, cons, dcons, dt, nsteps, nthreads=None):
    """nsteps fixed-dt RK3 steps of the state and of its tangent dcons."""
    c, dc = _f64(cons).copy(), _f64(dcons).copy()
    assert c.shape == dc.shape
    rc = lib().oracle_dual_rk3(ctypes.byref(grid), _p(c), _p(dc), float(dt), int(nsteps),
                               nthreads or nthreads_default())
    assert rc == 0
    return c, dc


def cfl_dt(grid, cons, cfl, nthreads=None):
    c = _f64(cons)
    return lib().oracle_cfl_dt(ctypes.byref(grid), _p(c), float(cfl), nthreads or nthreads_default())


def inter5(scheme, q, gamma=GAMMA_14):
    lib().oracle_set_gas(*gamma)
    q = _f64(q)
    return lib().oracle_inter5(scheme, _p(q))


def cons_to_prim(dim, c, gamma=GAMMA_14):
    lib().oracle_set_gas(*gamma)
    c = _f64(c)
    p = np.empty(dim + 2)
    lib().oracle_cons_to_prim(dim, _p(c), _p(p))
    return p


def eigen(dim, pl, pr, ax, gamma=GAMMA_14):
    lib().oracle_set_gas(*gamma)
    nv = dim + 2
    L = np.empty((nv, nv))
    R = np.empty((nv, nv))
    lib().oracle_eigen(dim, _p(_f64(pl)), _p(_f64(pr)), ax, _p(L), _p(R))
    return L, R


def prim_char_roundtrip(dim, pl, pr, ax, prim, gamma=GAMMA_14):
    lib().oracle_set_gas(*gamma)
    nv = dim + 2
    ch = np.empty(nv)
    back = np.empty(nv)
    lib().oracle_prim_char_roundtrip(dim, _p(_f64(pl)), _p(_f64(pr)), ax, _p(_f64(prim)), _p(ch), _p(back))
    return ch, back


def riemann(dim, flux, L, R, ax, gamma=GAMMA_14):
    lib().oracle_set_gas(*gamma)
    out = np.empty(dim + 2)
    lib().oracle_riemann(dim, flux, _p(_f64(L)), _p(_f64(R)), ax, _p(out))
    return out


def node_flux(dim, p, ax, gamma=GAMMA_14):
    lib().oracle_set_gas(*gamma)
    out = np.empty(dim + 2)
    lib().oracle_node_flux(dim, _p(_f64(p)), ax, _p(out))
    return out


def recon(dim, scheme, charwise, window, ax, gamma=GAMMA_14):
    lib().oracle_set_gas(*gamma)
    nv = dim + 2
    w = _f64(window).reshape(6 * nv)
    L = np.empty(nv)
    R = np.empty(nv)
    lib().oracle_recon(dim, scheme, int(charwise), _p(w), ax, _p(L), _p(R))
    return L, R


def accuracy_rhs(scheme, cells, ghost_l, ghost_r, h, gamma=GAMMA_14):
    lib().oracle_set_gas(*gamma)
    cells = _f64(cells)
    out = np.empty_like(cells)
    lib().oracle_accuracy_rhs(scheme, len(cells), _p(cells), _p(_f64(ghost_l)), _p(_f64(ghost_r)), float(h), _p(out))
    return out
