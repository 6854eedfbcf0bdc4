"""ctypes binding of ``libcongfd.so`` (C ABI: ``include/congfd.h``).

``Solver`` mirrors the reference's ``BlockSolver`` (zone_FD/src/src/blockSolver.cpp)
for the right-hand side and SSP-RK3 path: it owns the device state, evaluates
``rhs = +sum_d dF_d/dx_d`` (``SpDistributor::rhsSolve``) and advances with
``RK3_SSP``.  Enum values follow the reference (zone_FD/src/include/macro.hpp).
"""
import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class InterMethod(enum.IntEnum):
    """InterMethod (zone_FD/src/include/macro.hpp:37-77), implemented subset."""
    WCNS5 = 2            # weno5_JSchen
    WCNSZ5 = 3           # weno5_Z
    WCNS5CONGZ = 8       # Teno5_CongZ (TENO-S)
    WCNS5CONGZCT4 = 9    # Teno5_CongZCT4
    WCNS5CONGZCT7 = 10   # Teno5_CongZCT7 in 2D/3D; Teno5_CongAA in 1D, as the reference wires it
    TCNS5 = 11           # Teno5_Z (TENO5)
    TCNS5CT4 = 12
    TCNS5CT7 = 13
    MUCSLIN5 = 15        # musclIn5 (minmod MUSCL on q[1..3])
    CONGTCNS5CT5 = 22    # congTcns5_ZCT5 (Teno5_Z's arithmetic)
    CONGTCNS5CT10 = 23   # congTcns5_ZCT10
    WHFTCNSN = 17        # whf_TcnsN (chi_k < 1e-10)
    WHFTCNSNA = 18       # whf_TcnsN_A
    WHFTCNSNAS = 19      # whf_TcnsN_AS
    WHFTCNSNS = 20       # whf_TcnsN_S
    WHFTCNSNLAD = 21     # whf_TcnsN_LAD
    WHFAITCNSNS = 24     # whf_ai_TcnsN_S_1
    WHFAITCNSNA = 25     # whf_ai_TcnsN_A_2
    WHFAITCNSNAS_1 = 26  # whf_ai_TcnsN_AS_1
    WHFAITCNSNAS_2 = 27  # whf_ai_TcnsN_AS_2
    WHFAITCNSNLADS = 28  # whf_ai_TcnsN_LADS_yuanbao
    WHFAITCNSNAZS = 29   # whf_ai_TcnsN_AZS_yuanbao
    WHFAITCNSNMYASF002_1 = 30    # whf_ai_TcnsN_myASF002_1
    WHFAITCNSNMYASF002_2 = 31    # whf_ai_TcnsN_myASF002_2
    WHFAITCNSNMYASF002_AI1 = 32  # whf_ai_TcnsN_myASF002_ai1
    WHFZYCTCNSNMYASF002_1 = 33  # whf_zyc_TcnsN_myASF002_1 (TENO-S, C_T adapted to eta_min)
    WHFCOMPARE = 34      # whf_TcnsN_COMPARE
    NEW_WHFTCNSA = 36    # new_whf_TCNS_A ("TENO-A")
    NEW_WHFTCNSASF203_NOS = 38  # whf_TCNS_AS_myF203_NoS (TENO-S, cubic C_T law)
    NEW_WHFTCNSASF202_NOS = 39  # whf_TCNS_AS_myF202_NoS (TENO-S, quadratic C_T law)


class FluxMethod(enum.IntEnum):
    """FluxMethod (macro.hpp:84)."""
    HLLC = 0
    ROE = 1


class Recon(enum.IntEnum):
    CHAR = 0        # characteristic-wise (the reference's Euler path)
    COMPONENT = 1   # component-wise on primitive variables


class BC(enum.IntEnum):
    """BndType values (macro.hpp:33-34) plus extensions."""
    PERIODIC = 1
    SYMMETRY1D = 2
    DIRICHLET = 3
    OUTFLOW = 7       # SUPERSONICOUTLET
    SYMMETRYX = 10
    SYMMETRYY = 11
    SYMMETRYZ = 101
    MIRROR_COPY = 102
    DMR_TOP = 8       # DoubleMachUp
    DMR_WALL = 103    # double-Mach y-low wiring: Dirichlet for x < 1/6, SYMMETRYY beyond


DIFF_MND6 = 3


class cfd_config(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("h", ctypes.c_double * 3),
        ("inter_method", ctypes.c_int),
        ("flux_method", ctypes.c_int),
        ("diff_method", ctypes.c_int),
        ("recon", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("bc_prim", (ctypes.c_double * 5) * 6),
        ("device", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("origin", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
    ]


PRECISION_FP64, PRECISION_FP32, PRECISION_FP64_FMA = 0, 1, 2


class cfd_zone(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("h", ctypes.c_double * 3),
        ("bc", ctypes.c_int * 6),
        ("bc_prim", (ctypes.c_double * 5) * 6),
    ]


BATCH_GRAPH, BATCH_PERSISTENT, BATCH_PER_ZONE = 0, 1, 2


# Every entry point declared in include/congfd.h.
EXPORTS = [
    "cfd_last_error", "cfd_version", "cfd_create", "cfd_destroy", "cfd_set_state", "cfd_get_state",
    "cfd_eval_rhs", "cfd_rk3_step", "cfd_rk3_steps_cfl", "cfd_rk3_steps_cfl_noalloc", "cfd_cfl_dt",
    "cfd_conserved_totals",
    "cfd_steps_to_output", "cfd_synchronize", "cfd_device_buffers", "cfd_time_sweeps", "cfd_copy_probe", "cfd_set_keep_rhs", "cfd_rhs_linf", "cfd_autotune", "cfd_launch_plan", "cfd_set_launch_plan",
    "cfd_batch_create", "cfd_batch_destroy", "cfd_batch_set_state", "cfd_batch_get_state", "cfd_batch_steps_cfl",
    "cfd_batch_times", "cfd_batch_autotune", "cfd_batch_synchronize",
    "cfd_rhs_vjp", "cfd_adjoint_forward", "cfd_adjoint_backward", "cfd_set_time", "cfd_get_time",
]


def library_path():
    """In-tree library; CONGFD_LIB may name another build of the same sources."""
    return os.environ.get("CONGFD_LIB") or os.path.join(_HERE, "libcongfd.so")


_lib = None


def load_library():
    """Loads the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(make -C cfd_inwsl_amd/csrc)")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    pd = ctypes.POINTER(ctypes.c_double)
    lib.cfd_last_error.restype = ctypes.c_char_p
    lib.cfd_version.restype = ctypes.c_char_p
    lib.cfd_create.argtypes = [ctypes.POINTER(cfd_config), ctypes.POINTER(P)]
    lib.cfd_destroy.argtypes = [P]
    lib.cfd_set_state.argtypes = [P, pd]
    lib.cfd_get_state.argtypes = [P, pd]
    lib.cfd_eval_rhs.argtypes = [P, pd]
    lib.cfd_rk3_step.argtypes = [P, ctypes.c_double]
    lib.cfd_rk3_steps_cfl.argtypes = [P, ctypes.c_double, ctypes.c_int, pd]
    lib.cfd_rk3_steps_cfl_noalloc.argtypes = [P, ctypes.c_double, ctypes.c_int]
    lib.cfd_cfl_dt.argtypes = [P, ctypes.c_double, pd]
    lib.cfd_conserved_totals.argtypes = [P, pd]
    lib.cfd_synchronize.argtypes = [P]
    lib.cfd_device_buffers.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P)]
    lib.cfd_time_sweeps.argtypes = [P, ctypes.c_int, pd]
    lib.cfd_steps_to_output.argtypes = [P, ctypes.c_double, ctypes.c_double, pd, ctypes.c_int, pd,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.cfd_copy_probe.argtypes = [P, ctypes.c_int]
    lib.cfd_set_keep_rhs.argtypes = [P, ctypes.c_int]
    lib.cfd_rhs_linf.argtypes = [P, ctypes.c_int, pd]
    pi = ctypes.POINTER(ctypes.c_int)
    lib.cfd_autotune.argtypes = [P, ctypes.c_int, pi, pi]
    lib.cfd_launch_plan.argtypes = [P, pi, pi]
    lib.cfd_set_launch_plan.argtypes = [P, pi, pi]
    lib.cfd_batch_create.argtypes = [ctypes.POINTER(cfd_config), ctypes.c_int, ctypes.POINTER(cfd_zone), ctypes.POINTER(P)]
    lib.cfd_batch_destroy.argtypes = [P]
    lib.cfd_batch_set_state.argtypes = [P, ctypes.c_int, pd]
    lib.cfd_batch_get_state.argtypes = [P, ctypes.c_int, pd]
    lib.cfd_batch_steps_cfl.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    lib.cfd_batch_times.argtypes = [P, pd]
    lib.cfd_batch_autotune.argtypes = [P, ctypes.c_int]
    lib.cfd_batch_synchronize.argtypes = [P]
    lib.cfd_rhs_vjp.argtypes = [P, pd, pd]
    lib.cfd_set_time.argtypes = [P, ctypes.c_double]
    lib.cfd_get_time.argtypes = [P, pd]
    lib.cfd_adjoint_forward.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    lib.cfd_adjoint_backward.argtypes = [P, pd, pd]
    _lib = lib
    return lib


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class CfdError(RuntimeError):
    pass


class Solver:
    """One structured block of 1D/2D/3D Euler on one GPU.

    Parameters mirror the reference ``Info`` fields used on this path:
    ``n`` = cells per direction (``Info::icMax``), ``h`` = spacing
    (``Info::geth``), ``inter_method`` / ``flux_method`` (``SolverType``
    selection), boundary codes per face (x-, x+, y-, y+, z-, z+).
    """

    def __init__(self, dim, n, h, inter_method=InterMethod.WCNS5, flux_method=FluxMethod.ROE,
                 recon=Recon.CHAR, bc=None, bc_prim=None, device=0, precision=PRECISION_FP64, origin=0.0,
                 gravity=(0.0, 0.0, 0.0)):
        self.lib = load_library()
        self.dim = dim
        self.nv = dim + 2
        nn = list(n) + [1] * (3 - len(n))
        hh = list(h) if hasattr(h, "__len__") else [h] * dim
        hh = hh + [1.0] * (3 - len(hh))
        self.n = tuple(int(x) for x in nn)
        self.ncell = self.n[0] * self.n[1] * self.n[2]
        cfg = cfd_config()
        cfg.dim = dim
        for i in range(3):
            cfg.n[i] = self.n[i]
            cfg.h[i] = float(hh[i])
        cfg.inter_method = int(inter_method)
        cfg.flux_method = int(flux_method)
        cfg.diff_method = DIFF_MND6
        cfg.recon = int(recon)
        bc = list(bc) if bc is not None else [BC.OUTFLOW] * 6
        bc = bc + [BC.OUTFLOW] * (6 - len(bc))
        for i in range(6):
            cfg.bc[i] = int(bc[i])
            for v in range(5):
                cfg.bc_prim[i][v] = float(bc_prim[i][v]) if bc_prim is not None else 0.0
        cfg.device = device
        cfg.precision = int(precision)
        cfg.origin = float(origin)
        for i in range(3):
            cfg.gravity[i] = float(gravity[i]) if i < len(gravity) else 0.0
        self.cfg = cfg
        handle = ctypes.c_void_p()
        self._check(self.lib.cfd_create(ctypes.byref(cfg), ctypes.byref(handle)))
        self.h = handle

    def _check(self, rc):
        if rc != 0:
            raise CfdError(f"libcongfd error {rc}: {self.lib.cfd_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.cfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state in the reference layout: AoS [cell][nv], x fastest
    def set_state(self, cons):
        c = np.ascontiguousarray(cons, dtype=np.float64).reshape(self.ncell, self.nv)
        self._check(self.lib.cfd_set_state(self.h, _pd(c)))

    def get_state(self):
        out = np.empty((self.ncell, self.nv))
        self._check(self.lib.cfd_get_state(self.h, _pd(out)))
        return out

    def rhs(self, download=True):
        if not download:
            self._check(self.lib.cfd_eval_rhs(self.h, None))
            return None
        out = np.empty((self.ncell, self.nv))
        self._check(self.lib.cfd_eval_rhs(self.h, _pd(out)))
        return out

    def rk3_step(self, dt):
        self._check(self.lib.cfd_rk3_step(self.h, float(dt)))

    def rk3_steps_cfl(self, cfl, nsteps, want_time=True):
        if want_time:
            t = ctypes.c_double(0)
            self._check(self.lib.cfd_rk3_steps_cfl(self.h, float(cfl), int(nsteps), ctypes.byref(t)))
            return t.value
        self._check(self.lib.cfd_rk3_steps_cfl(self.h, float(cfl), int(nsteps), None))
        return None

    def set_keep_rhs(self, keep=True):
        self._check(self.lib.cfd_set_keep_rhs(self.h, int(bool(keep))))

    def rhs_linf(self, var=0):
        d = ctypes.c_double(0)
        self._check(self.lib.cfd_rhs_linf(self.h, int(var), ctypes.byref(d)))
        return d.value

    def cfl_dt(self, cfl):
        d = ctypes.c_double(0)
        self._check(self.lib.cfd_cfl_dt(self.h, float(cfl), ctypes.byref(d)))
        return d.value

    def totals(self):
        out = np.empty(self.nv)
        self._check(self.lib.cfd_conserved_totals(self.h, _pd(out)))
        return out

    def synchronize(self):
        self._check(self.lib.cfd_synchronize(self.h))

    def time_sweeps(self, reps=5):
        """Mean duration (ms) of each direction sweep of one RK stage (HIP events)."""
        out = np.empty(self.dim)
        self._check(self.lib.cfd_time_sweeps(self.h, int(reps), _pd(out)))
        return out

    def copy_probe(self, reps=1):
        """Plain device copy of the state planes (calibration of byte counters)."""
        self._check(self.lib.cfd_copy_probe(self.h, int(reps)))

    def autotune(self, reps=3):
        """Times the launch candidates of every direction; returns (tiles, segments)."""
        t = (ctypes.c_int * 3)()
        g = (ctypes.c_int * 3)()
        self._check(self.lib.cfd_autotune(self.h, int(reps), t, g))
        return list(t)[: self.dim], list(g)[: self.dim]

    def launch_plan(self):
        t = (ctypes.c_int * 3)()
        g = (ctypes.c_int * 3)()
        self._check(self.lib.cfd_launch_plan(self.h, t, g))
        return list(t)[: self.dim], list(g)[: self.dim]

    def set_launch_plan(self, tiles, segs):
        t = (ctypes.c_int * 3)(*(list(tiles) + [1] * (3 - len(tiles))))
        g = (ctypes.c_int * 3)(*(list(segs) + [1] * (3 - len(segs))))
        self._check(self.lib.cfd_set_launch_plan(self.h, t, g))

    def steps_to_output(self, cfl, output_dt, output_t=0.0, max_steps=64):
        """stepsLoopCFL on the device up to the next output time (or max_steps):
        returns (log [(t, dt, rhoLinf) per step], output_t, reached)."""
        ot = ctypes.c_double(output_t)
        log = np.empty((max_steps, 3))
        n, reached = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.cfd_steps_to_output(self.h, float(cfl), float(output_dt), ctypes.byref(ot),
                                                 int(max_steps), _pd(log), ctypes.byref(n), ctypes.byref(reached)))
        return log[:n.value].copy(), ot.value, bool(reached.value)

    def set_time(self, t):
        self._check(self.lib.cfd_set_time(self.h, float(t)))

    def get_time(self):
        t = ctypes.c_double(0)
        self._check(self.lib.cfd_get_time(self.h, ctypes.byref(t)))
        return t.value

    # ---- discrete adjoint (2D FP64, char-wise WCNS5 + HLLC) ----
    def rhs_vjp(self, cot):
        """(d rhs / d U)^T cot at the current state; cot and result AoS [cell][nv]."""
        cot = np.ascontiguousarray(cot, dtype=np.float64).reshape(self.ncell, self.nv)
        out = np.empty((self.ncell, self.nv))
        self._check(self.lib.cfd_rhs_vjp(self.h, _pd(cot), _pd(out)))
        return out

    def adjoint_forward(self, dt, nsteps, checkpoint_every=20):
        """nsteps fixed-dt RK3 steps, keeping a checkpoint every checkpoint_every
        steps for adjoint_backward (0: keep every stage state, no recomputation)."""
        self._check(self.lib.cfd_adjoint_forward(self.h, float(dt), int(nsteps), int(checkpoint_every)))

    def adjoint_backward(self, cot_final):
        """(d U_end / d U_0)^T cot_final for the last adjoint_forward."""
        cot = np.ascontiguousarray(cot_final, dtype=np.float64).reshape(self.ncell, self.nv)
        out = np.empty((self.ncell, self.nv))
        self._check(self.lib.cfd_adjoint_backward(self.h, _pd(cot), _pd(out)))
        return out

    def device_buffers(self):
        s, r = ctypes.c_void_p(), ctypes.c_void_p()
        self._check(self.lib.cfd_device_buffers(self.h, ctypes.byref(s), ctypes.byref(r)))
        return s.value, r.value


class Batch:
    """Independent zones of uneven sizes advanced together (cfd_batch_*)."""

    def __init__(self, dim, zones, inter_method=InterMethod.WCNS5, flux_method=FluxMethod.HLLC,
                 recon=Recon.CHAR, device=0, precision=PRECISION_FP64):
        """zones: list of dicts with keys n (tuple), h (float or tuple), bc (6 codes),
        optional bc_prim.  precision: PRECISION_FP64 or PRECISION_FP32 (FP32
        sweeps, FP64 state and stage combination)."""
        self.lib = load_library()
        self.dim = dim
        self.nv = dim + 2
        cfg = cfd_config()
        cfg.dim = dim
        cfg.inter_method = int(inter_method)
        cfg.flux_method = int(flux_method)
        cfg.diff_method = DIFF_MND6
        cfg.recon = int(recon)
        cfg.device = device
        cfg.precision = int(precision)
        arr = (cfd_zone * len(zones))()
        self.ncell = []
        for i, z in enumerate(zones):
            n = list(z["n"]) + [1] * (3 - len(z["n"]))
            h = z["h"] if hasattr(z["h"], "__len__") else [z["h"]] * dim
            h = list(h) + [1.0] * (3 - len(h))
            for d in range(3):
                arr[i].n[d] = int(n[d])
                arr[i].h[d] = float(h[d])
            bc = list(z.get("bc", [BC.OUTFLOW] * 6)) + [BC.OUTFLOW] * 6
            for f in range(6):
                arr[i].bc[f] = int(bc[f])
                for v in range(5):
                    arr[i].bc_prim[f][v] = float(z["bc_prim"][f][v]) if "bc_prim" in z else 0.0
            self.ncell.append(n[0] * n[1] * n[2])
        handle = ctypes.c_void_p()
        self._check(self.lib.cfd_batch_create(ctypes.byref(cfg), len(zones), arr, ctypes.byref(handle)))
        self.h = handle
        self.nz = len(zones)

    def _check(self, rc):
        if rc != 0:
            raise CfdError(f"libcongfd error {rc}: {self.lib.cfd_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.cfd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, zone, cons):
        c = np.ascontiguousarray(cons, dtype=np.float64).reshape(self.ncell[zone], self.nv)
        self._check(self.lib.cfd_batch_set_state(self.h, int(zone), _pd(c)))

    def get_state(self, zone):
        out = np.empty((self.ncell[zone], self.nv))
        self._check(self.lib.cfd_batch_get_state(self.h, int(zone), _pd(out)))
        return out

    def steps_cfl(self, cfl, nsteps, mode=BATCH_GRAPH):
        self._check(self.lib.cfd_batch_steps_cfl(self.h, float(cfl), int(nsteps), int(mode)))

    def times(self):
        out = np.empty(self.nz)
        self._check(self.lib.cfd_batch_times(self.h, _pd(out)))
        return out

    def autotune(self, reps=2):
        self._check(self.lib.cfd_batch_autotune(self.h, int(reps)))

    def synchronize(self):
        self._check(self.lib.cfd_batch_synchronize(self.h))
