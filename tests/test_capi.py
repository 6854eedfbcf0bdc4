"""CPU checks of the C-ABI library: it loads, exports every symbol declared in
include/congfd.h, and rejects bad configurations without touching a GPU."""
import ctypes
import os
import re

import pytest

from cfd_inwsl_amd import solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "congfd.h")).read()
    return sorted(set(re.findall(r"\b(cfd_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_match_binding():
    assert _declared() == sorted(solver.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = solver.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_config_struct_matches_header_size():
    # dim + n[3] | h[3] | 4 ints | bc[6] | bc_prim[6][5] | device, precision | origin | gravity[3]
    c = solver.cfd_config()
    assert ctypes.sizeof(c) == 4 + 12 + 24 + 16 + 24 + 240 + 4 + 4 + 8 + 24


def test_version_string():
    lib = solver.load_library()
    assert lib.cfd_version().decode().startswith("congfd-mi355x")


@pytest.mark.parametrize("field,value,msg", [
    ("dim", 4, "dim"),
    ("diff_method", 1, "MND6"),
    ("flux_method", 5, "flux_method"),
    ("inter_method", -1, "not implemented"),
    ("inter_method", 40, "not implemented"),  # beyond the reference's InterMethod values
])
def test_bad_config_is_rejected_before_any_gpu_call(field, value, msg):
    lib = solver.load_library()
    cfg = solver.cfd_config()
    cfg.dim = 2
    cfg.n[0], cfg.n[1], cfg.n[2] = 8, 8, 1
    cfg.h[0] = cfg.h[1] = cfg.h[2] = 0.1
    cfg.inter_method = 2
    cfg.flux_method = 0
    cfg.diff_method = 3
    cfg.recon = 0
    for i in range(6):
        cfg.bc[i] = 7
    setattr(cfg, field, value)
    h = ctypes.c_void_p()
    rc = lib.cfd_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc < 0
    assert msg in lib.cfd_last_error().decode()
    assert not h.value


def test_precision_codes_checked():
    """Unknown precision codes are refused by cfd_create, and batches refuse
    the FMA mode (they run FP64 or FP32), before any GPU call."""
    lib = solver.load_library()
    cfg = solver.cfd_config()
    cfg.dim = 1
    cfg.n[0], cfg.n[1], cfg.n[2] = 16, 1, 1
    cfg.h[0] = cfg.h[1] = cfg.h[2] = 0.1
    cfg.inter_method, cfg.flux_method, cfg.diff_method, cfg.recon = 2, 0, 3, 0
    for i in range(6):
        cfg.bc[i] = 7
    h = ctypes.c_void_p()
    cfg.precision = 7
    assert lib.cfd_create(ctypes.byref(cfg), ctypes.byref(h)) < 0
    assert "precision must be" in lib.cfd_last_error().decode()
    zones = (solver.cfd_zone * 1)()
    zones[0].n[0], zones[0].n[1], zones[0].n[2] = 16, 1, 1
    zones[0].h[0] = zones[0].h[1] = zones[0].h[2] = 0.1
    for i in range(6):
        zones[0].bc[i] = 7
    b = ctypes.c_void_p()
    for prec in (solver.PRECISION_FP64_FMA, 7):
        cfg.precision = prec
        assert lib.cfd_batch_create(ctypes.byref(cfg), 1, zones, ctypes.byref(b)) < 0
        assert "batches run in" in lib.cfd_last_error().decode()
        assert not b.value


def test_fma_precision_limited_to_its_schemes():
    """CFD_PRECISION_FP64_FMA is built for WENO-JS, WENO-Z, TENO5, TENO-S;
    any other scheme is refused before any GPU call."""
    lib = solver.load_library()
    cfg = solver.cfd_config()
    cfg.dim = 2
    cfg.n[0], cfg.n[1], cfg.n[2] = 16, 16, 1
    cfg.h[0] = cfg.h[1] = cfg.h[2] = 0.1
    cfg.flux_method, cfg.diff_method, cfg.recon = 0, 3, 0
    for i in range(6):
        cfg.bc[i] = 7
    cfg.precision = solver.PRECISION_FP64_FMA
    h = ctypes.c_void_p()
    for inter in (solver.InterMethod.NEW_WHFTCNSA, solver.InterMethod.MUCSLIN5, solver.InterMethod.WHFTCNSN):
        cfg.inter_method = int(inter)
        assert lib.cfd_create(ctypes.byref(cfg), ctypes.byref(h)) < 0
        assert "FP64_FMA is built for" in lib.cfd_last_error().decode()
        assert not h.value


def _batch_cfg():
    cfg = solver.cfd_config()
    cfg.dim = 2
    cfg.inter_method, cfg.flux_method, cfg.diff_method, cfg.recon = 2, 0, 3, 0
    return cfg


def _zones(specs):
    arr = (solver.cfd_zone * len(specs))()
    for i, (n, bc) in enumerate(specs):
        arr[i].n[0], arr[i].n[1], arr[i].n[2] = n[0], n[1], 1
        arr[i].h[0] = arr[i].h[1] = arr[i].h[2] = 0.1
        for f in range(6):
            arr[i].bc[f] = bc[f]
    return arr


@pytest.mark.parametrize("specs,msg", [
    # a mirroring face on an axis of fewer than 4 cells (its ghosts would read outside the zone)
    ([((16, 16), [7] * 6), ((3, 16), [2, 7, 7, 7, 7, 7])], "zone 1: mirror boundaries need at least 4 cells"),
    ([((16, 16), [7, 7, 102, 7, 7, 7]), ((16, 2), [7, 7, 7, 11, 7, 7])], "zone 1: mirror boundaries"),
    ([((16, 16), [7, 7, 7, 99, 7, 7])], "zone 0: unknown boundary code 99"),
    ([((16, 16), [8, 7, 7, 7, 7, 7])], "double-Mach boundaries belong on the y faces"),
])
def test_batch_rejects_bad_faces_before_any_gpu_call(specs, msg):
    lib = solver.load_library()
    cfg = _batch_cfg()
    arr = _zones(specs)
    b = ctypes.c_void_p()
    rc = lib.cfd_batch_create(ctypes.byref(cfg), len(specs), arr, ctypes.byref(b))
    assert rc == -1
    assert msg in lib.cfd_last_error().decode()
    assert not b.value


def test_batch_rejects_bad_recon():
    lib = solver.load_library()
    cfg = _batch_cfg()
    cfg.recon = 5
    b = ctypes.c_void_p()
    assert lib.cfd_batch_create(ctypes.byref(cfg), 1, _zones([((16, 16), [7] * 6)]), ctypes.byref(b)) == -1
    assert "recon" in lib.cfd_last_error().decode()
