"""Batched independent zones (config B) and launch-plan independence.

Every zone of a batch must evolve exactly as it does alone (same bits), in
all three execution modes (persistent work-queue launches replayed from a
hipGraph, eager persistent launches, single-zone kernels zone after zone), and
a change of tile shape or segment length must not change any result."""
import numpy as np
import pytest

from cfd_inwsl_amd import (BATCH_GRAPH, BATCH_PER_ZONE, BATCH_PERSISTENT, BC, PRECISION_FP32, PRECISION_FP64, Batch,
                           FluxMethod, InterMethod, Recon, Solver, cases)
from oracle import oracle as O
from tests.helpers import oracle_grid, random_state

pytestmark = pytest.mark.gpu

SIZES = [(40, 33), (70, 20), (25, 60), (128, 16), (33, 33), (17, 90)]


def _zones():
    out = []
    for i, (nx, ny) in enumerate(SIZES):
        cons, h = cases.riemann2d_quadrants(nx, ny, seed=100 + i)
        out.append(({"n": (nx, ny), "h": h, "bc": [BC.OUTFLOW] * 6}, cons))
    return out


def test_batch_modes_equal_single_zone_runs():
    zs = _zones()
    results = {}
    for mode in (BATCH_GRAPH, BATCH_PERSISTENT, BATCH_PER_ZONE):
        b = Batch(2, [z for z, _ in zs], InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR)
        for i, (_, c) in enumerate(zs):
            b.set_state(i, c)
        b.steps_cfl(0.5, 3, mode)
        results[mode] = ([b.get_state(i) for i in range(len(zs))], b.times())
        b.close()
    for i, (z, c) in enumerate(zs):
        s = Solver(2, z["n"], z["h"], InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, z["bc"])
        s.set_state(c)
        t = s.rk3_steps_cfl(0.5, 3)
        single = s.get_state()
        s.close()
        for mode, (states, times) in results.items():
            assert np.array_equal(states[i], single), (mode, i)
            assert times[i] == t, (mode, i)
    # one zone against the oracle
    z, c = zs[2]
    ref, tref = O.rk3(oracle_grid(2, z["n"], z["h"], InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, z["bc"]),
                      c, cfl=0.5, nsteps=3)
    assert np.array_equal(results[BATCH_GRAPH][0][2], ref)


def test_batch_graph_replays_match_eager():
    """The captured step replayed several times, with eager work between the
    replays, equals the eager persistent path after every replay."""
    zs = _zones()[:4]
    bg = Batch(2, [z for z, _ in zs], InterMethod.WCNS5, FluxMethod.ROE, Recon.CHAR)
    be = Batch(2, [z for z, _ in zs], InterMethod.WCNS5, FluxMethod.ROE, Recon.CHAR)
    for i, (_, c) in enumerate(zs):
        bg.set_state(i, c)
        be.set_state(i, c)
    for rep in range(3):
        bg.steps_cfl(0.5, 1, BATCH_GRAPH)
        be.steps_cfl(0.5, 1, BATCH_PERSISTENT)
        for i in range(len(zs)):
            assert np.array_equal(bg.get_state(i), be.get_state(i)), (rep, i)
    bg.close()
    be.close()


def test_batch_autotune_keeps_results():
    zs = _zones()
    b = Batch(2, [z for z, _ in zs], InterMethod.WCNSZ5, FluxMethod.HLLC, Recon.CHAR)
    ref = Batch(2, [z for z, _ in zs], InterMethod.WCNSZ5, FluxMethod.HLLC, Recon.CHAR)
    for i, (_, c) in enumerate(zs):
        b.set_state(i, c)
        ref.set_state(i, c)
    b.autotune(1)
    b.steps_cfl(0.5, 2, BATCH_GRAPH)
    ref.steps_cfl(0.5, 2, BATCH_PERSISTENT)
    for i in range(len(zs)):
        assert np.array_equal(b.get_state(i), ref.get_state(i)), i
    b.close()
    ref.close()


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_launch_plan_does_not_change_results(dim):
    n = {1: (300,), 2: (150, 70), 3: (40, 36, 30)}[dim]
    cons = random_state(dim, n, seed=11)
    s = Solver(dim, n, 0.05, InterMethod.TCNS5, FluxMethod.HLLC, Recon.CHAR, [BC.PERIODIC] * 6)
    s.set_state(cons)
    r0 = s.rhs()
    plans = [([0, 1, 1][:dim], [n[0], n[1] if dim > 1 else 1, n[2] if dim > 2 else 1][:dim]),
             ([1, 2, 2][:dim], [29, 13, 21][:dim]),
             ([0, 1, 2][:dim], [61, 45, 8][:dim]),
             ([1, 1, 1][:dim], [5, 7, 3][:dim])]
    for tiles, segs in plans:
        segs = [min(g, n[d]) for d, g in enumerate(segs)]
        s.set_launch_plan(tiles, segs)
        assert np.array_equal(s.rhs(), r0), (tiles, segs)
    tuned = s.autotune(2)
    r1 = s.rhs()
    s.close()
    print("autotuned plan", tuned, flush=True)
    assert np.array_equal(r0, r1)


def test_batch_research_constants_and_whf_refusal():
    """Batches carry the research-variant constants per zone: an adaptive
    TENO-S of the whf_ai family (C_T' from the per-zone table) evolves each
    zone as a single-zone run does.  The whf_TcnsN-type schemes, which have
    single-zone kernels only, are refused."""
    zs = _zones()[:3]
    b = Batch(2, [z for z, _ in zs], InterMethod.WHFAITCNSNAS_1, FluxMethod.HLLC, Recon.CHAR)
    for i, (_, c) in enumerate(zs):
        b.set_state(i, c)
    b.steps_cfl(0.5, 2, BATCH_GRAPH)
    got = [b.get_state(i) for i in range(len(zs))]
    b.close()
    for i, (z, c) in enumerate(zs):
        s = Solver(2, z["n"], z["h"], InterMethod.WHFAITCNSNAS_1, FluxMethod.HLLC, Recon.CHAR, z["bc"])
        s.set_state(c)
        s.rk3_steps_cfl(0.5, 2)
        assert np.array_equal(got[i], s.get_state()), i
        s.close()
    with pytest.raises(Exception, match="single zones only"):
        Batch(2, [z for z, _ in zs], InterMethod.WHFTCNSNLAD, FluxMethod.HLLC, Recon.CHAR)


def _config_b_zones(nzones, lo, hi, seed):
    """Config B's generator (the benchmark's): per-side sizes uniform in
    [lo, hi], seeded four-quadrant states, outflow faces."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=(nzones, 2))
    out = []
    for i, (nx, ny) in enumerate(sizes):
        cons, h = cases.riemann2d_quadrants(int(nx), int(ny), seed=seed + i)
        out.append(({"n": (int(nx), int(ny)), "h": h, "bc": [BC.OUTFLOW] * 6}, cons))
    return out


def _single_zone_run(z, c, steps, precision=PRECISION_FP64):
    s = Solver(2, z["n"], z["h"], InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR, z["bc"], precision=precision)
    s.set_state(c)
    t = s.rk3_steps_cfl(0.5, steps)
    out = s.get_state()
    s.close()
    return out, t


def test_config_b_wcns5_hllc_uneven_zones_128_512():
    """Config B's scheme (WCNS5 char-wise + HLLC) batched over 12 uneven
    four-quadrant zones with sides in [128, 512]: every zone equals its
    single-zone run bit for bit (state and clock), after an autotune, in the
    graph-replayed persistent mode; the two smallest zones equal the oracle."""
    zs = _config_b_zones(12, 128, 512, seed=4242)
    print("config B zones", [z["n"] for z, _ in zs], flush=True)
    steps = 2
    b = Batch(2, [z for z, _ in zs], InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR)
    for i, (_, c) in enumerate(zs):
        b.set_state(i, c)
    b.autotune(1)
    b.steps_cfl(0.5, steps, BATCH_GRAPH)
    got = [b.get_state(i) for i in range(len(zs))]
    times = b.times()
    b.close()
    for i, (z, c) in enumerate(zs):
        single, t = _single_zone_run(z, c, steps)
        assert np.array_equal(got[i], single), i
        assert times[i] == t, (i, times[i], t)
    small = sorted(range(len(zs)), key=lambda i: zs[i][0]["n"][0] * zs[i][0]["n"][1])[:2]
    for i in small:
        z, c = zs[i]
        ref, tref = O.rk3(oracle_grid(2, z["n"], z["h"], InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR, z["bc"]),
                          c, cfl=0.5, nsteps=steps)
        print(f"config B zone {i} {z['n']} vs oracle: bitwise={np.array_equal(got[i], ref)}", flush=True)
        assert np.array_equal(got[i], ref), i
        assert times[i] == tref, i


def test_config_b_96_zones_128_1024():
    """Config B at its stated size: 96 zones, sides uniform in [128, 1024]
    (the benchmark's workload, about 30 M cells), two graph-replayed steps;
    a sample of zones (the largest, the smallest and six more) equals its
    single-zone run bit for bit."""
    zs = _config_b_zones(96, 128, 1024, seed=20261203)
    steps = 2
    b = Batch(2, [z for z, _ in zs], InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR)
    for i, (_, c) in enumerate(zs):
        b.set_state(i, c)
    b.steps_cfl(0.5, steps, BATCH_GRAPH)
    cells = [z["n"][0] * z["n"][1] for z, _ in zs]
    sample = sorted({int(np.argmax(cells)), int(np.argmin(cells)), 3, 17, 40, 58, 77, 95})
    times = b.times()
    for i in sample:
        z, c = zs[i]
        single, t = _single_zone_run(z, c, steps)
        got = b.get_state(i)
        print(f"96-zone batch, zone {i} {z['n']}: bitwise={np.array_equal(got, single)}", flush=True)
        assert np.array_equal(got, single), i
        assert times[i] == t, i
    b.close()


def test_batch_fp32_equals_single_zone_fp32_runs():
    """FP32 batches (FP32 sweeps over FP32 pools, FP64 state and stage
    combination): every zone equals its FP32 single-zone run bit for bit in
    all three modes (same kernels, same per-zone dt), and the batch's
    per-variable L1 / Linf difference to the FP64 batch is bounded."""
    zs = _config_b_zones(6, 48, 160, seed=777)
    steps = 3
    res = {}
    for mode in (BATCH_GRAPH, BATCH_PERSISTENT, BATCH_PER_ZONE):
        b = Batch(2, [z for z, _ in zs], InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR, precision=PRECISION_FP32)
        for i, (_, c) in enumerate(zs):
            b.set_state(i, c)
        b.steps_cfl(0.5, steps, mode)
        res[mode] = ([b.get_state(i) for i in range(len(zs))], b.times())
        b.close()
    b64 = Batch(2, [z for z, _ in zs], InterMethod.WCNS5, FluxMethod.HLLC, Recon.CHAR)
    for i, (_, c) in enumerate(zs):
        b64.set_state(i, c)
    b64.steps_cfl(0.5, steps, BATCH_GRAPH)
    f64 = [b64.get_state(i) for i in range(len(zs))]
    b64.close()
    for i, (z, c) in enumerate(zs):
        single, t = _single_zone_run(z, c, steps, precision=PRECISION_FP32)
        for mode, (states, times) in res.items():
            assert np.array_equal(states[i], single), (mode, i)
            assert times[i] == t, (mode, i)
        scale = np.abs(f64[i]).max(axis=0)
        l1 = np.abs(single - f64[i]).mean(axis=0) / scale
        linf = np.abs(single - f64[i]).max(axis=0) / scale
        print(f"fp32 batch zone {i} {z['n']} vs fp64: L1 {l1} Linf {linf}", flush=True)
        assert np.all(l1 < 1e-5) and np.all(linf < 1e-3), i
