"""wavesim's per-step statistics (crates/wavesim/src/lib.rs:849-937, 1099-1160, 1276-1300) on the
oracle, pinned by closed forms; StepWatch through the library's host-only entry points."""

from __future__ import annotations

import math

import numpy as np

from oracle import oracle as orc
from wavesim_amd.grid import Bathymetry, Grid, State


def oscillating(grid, bed_depth, t, mean, a, b, omega):
    s = State.zeros(grid)
    grid.inner(s.h)[:, :] = bed_depth + mean + a * math.cos(omega * t) + b * math.sin(omega * t)
    s.fill_walls(grid)
    s.time = t
    return s


def test_first_harmonic_and_wave_heights_of_a_pure_oscillation():
    grid = Grid(8, 6, 2.0, 2.0)
    bed = Bathymetry.flat(grid, 5.0)
    o = orc.OracleSolver(grid, bed)
    period, mean, a, b = 12.0, 0.3, 0.7, -0.4
    omega = 2 * math.pi / period
    st = orc.OracleStats(o, 0.0, period)
    dt = 0.1
    for n in range(1, 361):  # three periods
        t = n * dt
        st.record(oscillating(grid, 5.0, t, mean, a, b, omega), dt)
    amp = st.harmonic()
    assert amp is not None
    fa, fb = amp
    # an exact model is fitted exactly by least squares, whatever the weights
    assert np.max(np.abs(fa - a)) < 1e-9 and np.max(np.abs(fb + b)) < 1e-9
    r = st.read()
    assert r["waves"] == 2  # closed at the ends of periods 1 and 2
    assert abs(r["seconds"] - 36.0) < 1e-9
    crest = math.hypot(a, b)
    assert np.all(r["wave_first"] <= 2 * crest + 1e-12)
    assert np.all(r["wave_first"] >= 2 * crest * math.cos(omega * dt))  # sampled every dt
    assert np.allclose(r["eta_max"], mean + crest, atol=1e-3)
    assert np.allclose(r["eta_min"], mean - crest, atol=1e-3)
    assert np.all(r["breaking_seconds"] == 0.0)
    # Stage::finish closes the third wave, which ended with the last step (t = 36 s)
    st.finish()
    r3 = st.read()
    assert r3["waves"] == 3 and np.array_equal(r3["wave_first"], r["wave_first"])
    assert np.allclose(r3["wave_last"], r["wave_last"], rtol=1e-3)
    st.finish()  # nothing more is due
    assert st.read()["waves"] == 3


def test_harmonic_is_none_before_half_a_period():
    grid = Grid(4, 4, 1.0, 1.0)
    bed = Bathymetry.flat(grid, 2.0)
    o = orc.OracleSolver(grid, bed)
    st = orc.OracleStats(o, 0.0, 10.0)
    for n in range(1, 4):
        st.record(oscillating(grid, 2.0, 0.5 * n, 0.0, 0.1, 0.0, 2 * math.pi / 10.0), 0.5)
    assert st.harmonic() is None


def test_recording_starts_at_from_and_counts_breaking_cells():
    grid = Grid(6, 4, 1.0, 1.0)
    bed = Bathymetry.flat(grid, 1.0)
    o = orc.OracleSolver(grid, bed)
    st = orc.OracleStats(o, 10.0, 8.0)
    s = State.zeros(grid)
    grid.inner(s.h)[:, :] = 1.0
    grid.inner(s.h)[:, 4:] = 3.0  # surface 2 m above the 1 m deep bed: eta / h = 0.67 -> breaking
    s.fill_walls(grid)
    s.time = 10.3
    st.record(s, 0.5)  # only 0.3 s of the step lies after `from`
    s.time = 10.8
    st.record(s, 0.5)
    r = st.read()
    inside = (10.3 - 10.0) + 0.5
    assert r["seconds"] == inside
    bs = r["breaking_seconds"].reshape(grid.ny, grid.nx)
    # columns 4-5 by the ratio eta / h, column 3 by the slope of the step (1 > 0.65)
    assert np.all(bs[:, 3:] == inside) and np.all(bs[:, :3] == 0.0)


def test_step_watch_reports_a_collapse_once_a_block_ends():
    from wavesim_amd.solver import StepWatch

    w = StepWatch(0.1)
    for _ in range(999):
        assert w.collapsed(0.01) is None
    assert abs(w.collapsed(0.01) - 10.0) < 1e-9
    for _ in range(1000):
        assert w.collapsed(0.05) is None  # a fifth of the still sea's step or more
    w = StepWatch(float("inf"))  # no water: the watch is off
    for _ in range(2000):
        assert w.collapsed(1e-9) is None
