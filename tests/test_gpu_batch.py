"""Many instances in one launch (wc_batch_*): every instance of a batch must end exactly where
it ends when stepped alone (the FP32 form with the one-kernel PCG is deterministic, and a
batched launch runs each instance's blocks with the single launch's arithmetic), whatever the
persistent launch widths; an instance whose step stops being finite stops alone."""

from __future__ import annotations

import numpy as np
import pytest

from wavesim_amd import configs
from wavesim_amd.grid import State
from wavesim_amd.solver import PRECISION_FP32, Batch, Solver

pytestmark = pytest.mark.gpu


def workloads():
    """Four reef-sheet instances of uneven size (seeded like the sweep, the smaller sizes) and a
    periodic rough bed."""
    out = []
    seed = 3000
    while len(out) < 4:
        w = configs.c3_reef_instance(seed)
        if w.grid.nx <= 768 and all(o.grid.nx != w.grid.nx or o.notes["lobes"] != w.notes["lobes"]
                                    for o in out):
            out.append(w)
        seed += 1
    out.append(configs.c1_rough_bed(nx=320, ny=96, seed=5))
    return out


def fresh(w, arrays=None, t0=0.0, walk=1):
    g = w.solver(Solver).with_precision(PRECISION_FP32)
    if walk != 1:
        g.set_option("pcg_tiles_per_block", walk)
    s = w.still_water() if arrays is None else State(*(a.copy() for a in arrays), t0)
    return g, s


def run_alone(ws, starts, steps):
    out = []
    for w, (arr, t0) in zip(ws, starts):
        g, s = fresh(w, arr, t0)
        n, last = g.run_steps(s, steps)
        out.append((n, s.time, [a.copy() for a in s.arrays()[:3]], g.dispersion_stats()))
    return out


def run_batched(ws, starts, steps, widths=None, graph=True):
    pairs = [fresh(w, arr, t0) for w, (arr, t0) in zip(ws, starts)]
    b = Batch([g for g, _ in pairs])
    b.set_option("graph", int(graph))
    if widths:
        b.set_option("grid_nt", widths[0]).set_option("grid_pt", widths[1])
    done, last = b.run_steps([s for _, s in pairs], steps)
    res = [(d, s.time, [a.copy() for a in s.arrays()[:3]], g.dispersion_stats())
           for d, (g, s) in zip(done, pairs)]
    return res, b.launch_widths(), last


def spun_up(ws, steps):
    starts = []
    for w in ws:
        g, s = fresh(w)
        g.run_steps(s, steps)
        starts.append(([a.copy() for a in s.arrays()], s.time))
    return starts


def assert_same(alone, batched):
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert a[0] == b[0], k
        assert a[1] == b[1], k
        for x, y in zip(a[2], b[2]):
            assert np.array_equal(x, y), k
        assert a[3] == b[3], k


@pytest.mark.parametrize("graph,walk", [(True, 1), (False, 1), (True, 4), (False, 4)])
def test_a_batch_steps_each_instance_as_it_steps_alone(graph, walk):
    """With the fixed stretches of each step replayed from HIP graphs (the default) and
    without, and with PCG blocks taking 4 tiles or more; compared after every 5 steps, so that
    every replay between two looks counts."""
    ws = workloads()
    starts = spun_up(ws, 120)
    pairs = [fresh(w, arr, t0, walk) for w, (arr, t0) in zip(ws, starts)]
    solo = [fresh(w, arr, t0, walk) for w, (arr, t0) in zip(ws, starts)]
    b = Batch([g for g, _ in pairs])
    b.set_option("graph", int(graph))
    for chunk in range(5):
        done, _ = b.run_steps([s for _, s in pairs], 5)
        assert done == [5] * len(ws)
        for k, ((g, s), (g1, s1)) in enumerate(zip(pairs, solo)):
            assert g1.run_steps(s1, 5)[0] == 5
            assert s.time == s1.time, (chunk, k)
            for x, y in zip(s.arrays()[:3], s1.arrays()[:3]):
                assert np.array_equal(x, y), (chunk, k)
            assert g.dispersion_stats() == g1.dispersion_stats(), (chunk, k)
    widths = b.launch_widths()
    print("instances", [(w.grid.nx, w.grid.ny) for w in ws], "widths", widths,
          "iterations", [g.dispersion_stats().iterations for g, _ in solo], flush=True)
    assert all(g.dispersion_stats().iterations > 0 for g, _ in solo)
    assert widths[0] > 0 and widths[1] > 0


@pytest.mark.parametrize("widths", [(1, 1), (37, 5), (100000, 100000)])
def test_the_launch_width_changes_nothing(widths):
    ws = workloads()[:3]
    starts = spun_up(ws, 60)
    alone = run_alone(ws, starts, 8)
    batched, got, _ = run_batched(ws, starts, 8, widths)
    assert got == widths
    assert_same(alone, batched)


@pytest.mark.parametrize("rows", [0, 4, 8, 16])
def test_the_setup_tile_height_changes_nothing(rows):
    """The batched SGN set-up's tile height (64 x rows cells a block): the tuner's choice (the
    fastest of 4, 8 and 16, timed at the first run) and each height fixed step every instance
    to the bits it reaches alone; other heights are refused."""
    ws = workloads()[:3]
    starts = spun_up(ws, 40)
    alone = run_alone(ws, starts, 6)
    pairs = [fresh(w, arr, t0) for w, (arr, t0) in zip(ws, starts)]
    b = Batch([g for g, _ in pairs])
    if rows:
        b.set_option("setup_rows", rows)
    with pytest.raises(ValueError):
        b.set_option("setup_rows", 12)
    done, _ = b.run_steps([s for _, s in pairs], 6)
    assert done == [6] * len(ws)
    used, tried = b.setup_rows()
    print(f"setup rows {rows}: in use {used}, timed {tried}", flush=True)
    if rows == 0:
        assert [r for r, _ in tried] == [4, 8, 16]
        assert used == min(tried, key=lambda t: t[1])[0]
    else:
        assert used == rows and tried == []
    batched = [(d, s.time, [a.copy() for a in s.arrays()[:3]], g.dispersion_stats())
               for d, (g, s) in zip(done, pairs)]
    assert_same(alone, batched)


@pytest.mark.parametrize("rows", [0, 8, 13, 100000, "tiles"])
def test_the_flux_band_height_changes_nothing(rows):
    """The batched marching flux's band height (each wave's tile is 58 x rows cells): the
    tuner's choice (timed at the first run, the fastest of its candidates), fixed heights that
    divide no grid or exceed every grid, and the 64x4 tile flux all step each instance to the
    bits it reaches alone."""
    ws = workloads()[:3]
    starts = spun_up(ws, 60)
    alone = run_alone(ws, starts, 8)
    pairs = [fresh(w, arr, t0) for w, (arr, t0) in zip(ws, starts)]
    if rows == "tiles":
        for g, _ in pairs:
            g.set_option("flux_kernel", 2)
    b = Batch([g for g, _ in pairs])
    if rows != "tiles" and rows:
        b.set_option("flux_rows", rows)
    done, _ = b.run_steps([s for _, s in pairs], 8)
    assert done == [8] * len(ws)
    used, tried = b.flux_rows()
    print(f"flux rows {rows}: in use {used}, timed {tried}", flush=True)
    if rows == "tiles":
        assert used == 0 and tried == []
    elif rows == 0:
        assert len(tried) >= 2 and used == min(tried, key=lambda t: t[1])[0]
    else:
        assert used == rows and tried == []
    batched = [(d, s.time, [a.copy() for a in s.arrays()[:3]], g.dispersion_stats())
               for d, (g, s) in zip(done, pairs)]
    assert_same(alone, batched)


def test_a_solo_run_between_batch_runs_changes_nothing():
    """Instances stepped alone with other PCG layouts between two batch runs (other lanes and
    blocks write lanes the batch relies on reading zero) still step in the batch as alone; a
    batch refuses an instance whose settings changed since it was made."""
    ws = workloads()[:3]
    starts = spun_up(ws, 40)
    pairs = [fresh(w, arr, t0) for w, (arr, t0) in zip(ws, starts)]
    solo = [fresh(w, arr, t0) for w, (arr, t0) in zip(ws, starts)]
    b = Batch([g for g, _ in pairs])
    assert b.run_steps([s for _, s in pairs], 3)[0] == [3] * 3
    for (g, s), (g1, s1) in zip(pairs, solo):
        g1.run_steps(s1, 3)
        # (setting pcg_lanes clears the lanes itself, so it comes first and the layouts that
        # leave other lanes written come last)
        for opts in ({"pcg_lanes": 512}, {"pcg_lanes": 256, "pcg_march": 1},
                     {"pcg_march": 0, "pcg_kernels": 3}):
            for k, v in opts.items():
                g.set_option(k, v)
                g1.set_option(k, v)
            g.run_steps(s, 2)
            g1.run_steps(s1, 2)
        g.set_option("pcg_kernels", 4)
        g1.set_option("pcg_kernels", 4)
    assert b.run_steps([s for _, s in pairs], 5)[0] == [5] * 3
    for (g, s), (g1, s1) in zip(pairs, solo):
        assert g1.run_steps(s1, 5)[0] == 5
        assert s.time == s1.time
        for x, y in zip(s.arrays()[:3], s1.arrays()[:3]):
            assert np.array_equal(x, y)
    pairs[1][0].set_option("pcg_lanes", 512)
    with pytest.raises(ValueError, match="settings changed"):
        b.run_steps([s for _, s in pairs], 1)


def test_an_instance_that_stops_stops_alone():
    ws = workloads()[:3]
    starts = spun_up(ws, 40)
    h = starts[1][0][0].copy()
    h[ws[1].grid.idx(20, 20)] = np.nan
    starts[1] = ([h] + starts[1][0][1:], starts[1][1])
    alone = run_alone([ws[0], ws[2]], [starts[0], starts[2]], 10)
    batched, _, last = run_batched(ws, starts, 10)
    assert batched[1][0] == 0 and np.isnan(last[1])
    assert_same(alone, [batched[0], batched[2]])
