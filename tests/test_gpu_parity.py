"""Parity of the GPU path with the CPU oracle, through the C ABI.

The oracle runs in the GPU's canonical summation order with the portable libm (the two things
the reference's text leaves to the platform; the oracle's other mode, the reference's row order
and glibc, is compared with this one in test_oracle_modes.py). In that mode the FP64 GPU path
must agree bit for bit: h, hu, hv and breaking on every padded cell, the time, every stable dt,
the PCG statistics.
"""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import known_answers as ka
from helpers import bore_state
from oracle import oracle as orc
from wavesim_amd import configs
from wavesim_amd.config import Dispersion, Order, Sponge, WaveMaker, linear_wave
from wavesim_amd.grid import GHOST, Bathymetry, Grid, State
from wavesim_amd.solver import Solver

pytestmark = pytest.mark.gpu


def oracle(grid, bed, order=orc.DOT_CANONICAL16):
    """The oracle in the GPU's summation order: 64x16 tiles for the default FP64 PCG, 64x4
    tiles for the PCG variants pcg_kernels 1-3."""
    return orc.OracleSolver(grid, bed, dot_order=order, math=orc.MATH_PORTABLE)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_float(a: float, b: float) -> bool:
    return (math.isnan(a) and math.isnan(b)) or a == b


def assert_states_equal(sg: State, so: State, label: str):
    ag, ao = sg.arrays(), so.arrays()
    for name, x, y in zip(("h", "hu", "hv", "breaking"), ag, ao):
        bad = np.flatnonzero(~((x == y) | (np.isnan(x) & np.isnan(y))))
        assert bad.size == 0, (
            f"{label}: {name} differs at {bad.size} cells, first {bad[:5]}: "
            f"{x[bad[:3]]} vs {y[bad[:3]]}")
    assert sg.time == so.time, (label, sg.time, so.time)


def lockstep(grid, bed, configure, state: State, steps: int, label: str, fixed_dt=None,
             options=None):
    g = Solver(grid, bed.copy())
    for k, v in (options or {}).items():
        g.set_option(k, v)
    g = configure(g)
    order = (orc.DOT_CANONICAL if (options or {}).get("pcg_kernels") in (1, 2, 3)
             else orc.DOT_CANONICAL16)
    o = configure(oracle(grid, bed.copy(), order))
    sg, so = state.clone(), state.clone()
    say(f"[{label}] {grid.nx}x{grid.ny}, {steps} steps")
    for n in range(steps):
        if fixed_dt is None:
            dg, do = g.stable_dt(sg), o.stable_dt(so)
            assert same_float(dg, do), (label, n, dg, do)
            if not (math.isfinite(dg) and dg > 0):
                break
        else:
            dg = do = fixed_dt
        g.step(sg, dg)
        o.step(so, do)
    assert_states_equal(sg, so, label)
    assert g.dispersion_stats() == o.dispersion_stats(), (g.dispersion_stats(),
                                                          o.dispersion_stats())
    return g, o, sg, so


def hump_state(grid: Grid, bed: Bathymetry, amp=0.6, x0=None, y0=None, r=6.0):
    s = State.lake_at_rest(grid, bed, 0.0)
    x, y = grid.centres()
    X, Y = np.meshgrid(x[GHOST:GHOST + grid.nx], y[GHOST:GHOST + grid.ny])
    x0 = grid.nx * grid.dx * 0.4 if x0 is None else x0
    y0 = grid.ny * grid.dy * 0.5 if y0 is None else y0
    h = grid.inner(s.h)
    h += np.where(h > 0, amp * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (2 * r * r)), 0.0)
    s.fill_walls(grid)
    return s


@pytest.mark.parametrize("flux", [1, 2, 3])
def test_plane_beach_fv_2000_steps_is_bitwise_the_oracle(flux):
    w = configs.c0_plane_beach()
    lockstep(w.grid, w.bed, w.configure, w.still_water(), w.steps, f"{w.name} flux={flux}",
             options={"flux_kernel": flux})


@pytest.mark.parametrize("flux", [1, 2, 3])
@pytest.mark.parametrize("order", [Order.First, Order.Second, Order.Fifth])
def test_each_order_over_a_rough_bed_with_dry_land(order, flux):
    grid = Grid(96, 40, 3.0, 3.0)
    bed = Bathymetry.rough(grid, 1.5, 2.5)
    lockstep(grid, bed, lambda s: s.with_order(order), hump_state(grid, bed, amp=1.0),
             150, f"rough-dry {order.name} flux={flux}", options={"flux_kernel": flux})


@pytest.mark.parametrize("pcg", [1, 2, 3, 5])
@pytest.mark.parametrize("periodic", [False, True])
def test_sgn_fifth_walls_and_periodic_seam(periodic, pcg):
    grid = Grid(40, 30, 1.0, 1.0)
    bed = Bathymetry.flat(grid, 2.0)

    def configure(s):
        s = s.with_order(Order.Fifth).with_dispersion(Dispersion())
        return s.with_periodic_y() if periodic else s

    s = State.lake_at_rest(grid, bed, 0.0)
    x, y = grid.centres()
    X, Y = np.meshgrid(x[GHOST:GHOST + grid.nx], y[GHOST:GHOST + grid.ny])
    cy = 0.5 if periodic else 2.5
    grid.inner(s.h)[:, :] += 1.3 * np.exp(-((X - 20.0) ** 2 + (Y - cy) ** 2) / 32.0)
    lockstep(grid, bed, configure, s, 120, f"sgn seam periodic={periodic} pcg={pcg}",
             options={"pcg_kernels": pcg})


@pytest.mark.parametrize("cap", [7, 8])
def test_sgn_solves_stopped_after_an_odd_or_an_even_iteration(cap):
    """Every solve stopped by max_iterations after 7 (odd) or 8 (even) iterations, on a grid of
    15 tiles: the FP64 PCG applies the solution's steps two iterations at a time and the last
    one alone after an odd count, and the solution still ends where the oracle's does."""
    grid = Grid(150, 70, 1.0, 1.0)
    bed = Bathymetry.flat(grid, 2.0)

    def configure(s):
        return s.with_order(Order.Fifth).with_dispersion(
            Dispersion(tolerance=1e-14, max_iterations=cap))

    s = State.lake_at_rest(grid, bed, 0.0)
    x, y = grid.centres()
    X, Y = np.meshgrid(x[GHOST:GHOST + grid.nx], y[GHOST:GHOST + grid.ny])
    grid.inner(s.h)[:, :] += 0.9 * np.exp(-((X - 60.0) ** 2 + (Y - 30.0) ** 2) / 40.0)
    g, _, _, _ = lockstep(grid, bed, configure, s, 40, f"sgn capped at {cap}")
    st = g.dispersion_stats()
    assert st.solves > 0 and st.unconverged == st.solves and st.iterations == cap * st.solves


@pytest.mark.parametrize("pcg", [1, 2, 3, 5])
def test_sgn_rough_bed_tight_tolerance_with_maker_sponges_and_friction(pcg):
    w = configs.c1_rough_bed(nx=160, ny=64, steps=200)

    def configure(s):
        s = w.configure(s)
        s = s.with_sponge(Sponge(6, 0.8).edges(False, True, True, True)).with_manning(0.025)
        return s

    # start mid-ramp so the maker and the dispersion have something to do
    s = w.still_water()
    s.time = 20.0
    lockstep(w.grid, w.bed, configure, s, w.steps, f"sgn rough bed pcg={pcg}",
             options={"pcg_kernels": pcg, "pcg_batch": 3, "flux_kernel": 1 + pcg % 3})


def test_large_reef_tile_is_bitwise_the_oracle():
    """A 256x96 tile of the 4096x2048 reef's bed (the slope from the wave-maker at the 15 m
    contour to the 2 m contour, with the reef lobes and the rough overlay), with its physics:
    fifth order, SGN to 1e-8, Manning, sponges, the oblique wave-maker started near full
    amplitude."""
    w = configs.c4_crop()
    s = w.still_water()
    s.time = 60.0
    lockstep(w.grid, w.bed, w.configure, s, 150, "large reef tile")


@pytest.mark.parametrize("order,window", [(3, 6), (2, 2)])
def test_fp64_extrapolated_warm_start_is_bitwise_the_oracle(order, window):
    """The extrapolated warm start (no counterpart in the reference: its solves start from the
    previous solution) as an option of the FP64 path: ring slots, weights and the start, then
    the reference's PCG from it, bit for bit the oracle's with_warm_start."""
    w = configs.c1_rough_bed(nx=160, ny=64, steps=120)
    s = w.still_water()
    s.time = 40.0

    def configure(sv):
        return w.configure(sv)

    g = Solver(w.grid, w.bed.copy())
    g = configure(g).set_option("fp64_warm_start", order).set_option("warm_start_window", window)
    o = configure(oracle(w.grid, w.bed.copy())).with_warm_start(order, window)
    sg, so = s.clone(), s.clone()
    for n in range(w.steps):
        dg, do = g.stable_dt(sg), o.stable_dt(so)
        assert same_float(dg, do), (n, dg, do)
        g.step(sg, dg)
        o.step(so, do)
    assert_states_equal(sg, so, f"fp64 warm start {order} {window}")
    assert g.dispersion_stats() == o.dispersion_stats()
    print(f"fp64 warm start {order} window {window}: {g.dispersion_stats()}", flush=True)


def test_sgn_second_order_linear_dispersion_and_held_cells():
    grid = Grid(120, 8, 1.0, 1.0)
    bed = Bathymetry.flat(grid, 2.0)
    s = State.lake_at_rest(grid, bed, 0.0)
    x, _ = grid.centres()
    grid.inner(s.h)[:, :] += 0.3 * np.exp(-(((x[GHOST:GHOST + grid.nx] - 40.0) / 5.0) ** 2))
    held = [i >= 80 for _ in range(grid.ny) for i in range(grid.nx)]

    def configure(sv):
        return (sv.with_order(Order.Second).with_dispersion(Dispersion(nonlinear=False))
                .with_shallow_cells(held))

    lockstep(grid, bed, configure, s, 200, "held cells", fixed_dt=0.05)


def test_oblique_maker_in_a_periodic_channel():
    grid = Grid(200, 24, 2.0, 2.5)
    bed = Bathymetry.from_fn(grid, lambda x, y: -6.0 + x / 100.0)
    ly = grid.ny * grid.dy
    period = 8.0
    maker_x = 40.0
    # the depth the wave-maker reads (solver.rs:405-406) sets k; one alongshore wave must fit
    i = round(maker_x / grid.dx - 0.5) + GHOST
    depth = -bed.b[grid.idx(i, GHOST + grid.ny // 2)]
    k = linear_wave(2.0 * math.pi / period, depth, False).k
    angle = math.asin(2.0 * math.pi / (k * ly))
    maker = WaveMaker(maker_x, 0.1, period, angle, 4.0, 1.0)

    def configure(sv):
        return (sv.with_periodic_y().with_wavemaker(maker)
                .with_sponge(Sponge(12, 1.0).edges(True, True, False, False))
                .with_sponge(Sponge(5, 0.3).edges(False, True, False, False)))

    lockstep(grid, bed, configure, State.lake_at_rest(grid, bed, 0.0), 300, "oblique periodic")


def test_diagnostics_match_the_oracle():
    grid = Grid(70, 33, 2.0, 3.0)
    bed = Bathymetry.rough(grid, 1.0, 2.0)
    s = hump_state(grid, bed, amp=0.8)
    g = Solver(grid, bed.copy()).with_order(Order.Fifth).with_dispersion(Dispersion())
    o = oracle(grid, bed.copy()).with_order(Order.Fifth).with_dispersion(Dispersion())
    sg, so = s.clone(), s.clone()
    for _ in range(20):
        g.step(sg, 0.1)
        o.step(so, 0.1)
    assert_states_equal(sg, so, "diagnostics")
    np.testing.assert_array_equal(g.breaking_indicator(sg), o.breaking_indicator(so))
    assert g.volume(sg) == o.volume(so)  # both in the canonical order
    assert abs(g.volume(sg) - sg.volume(grid)) <= 1e-13 * sg.volume(grid)  # vs serial order
    assert g.max_abs_momentum(sg) == sg.max_abs_momentum(grid)
    assert g.min_depth(sg) == sg.min_depth(grid)
    assert g.limiting_cell(sg) == o.limiting_cell(so)


def test_state_round_trips_and_stays_on_the_device():
    grid = Grid(33, 17, 1.0, 1.0)
    bed = Bathymetry.flat(grid, 1.0)
    s = hump_state(grid, bed, amp=0.2, r=3.0)
    before = [a.copy() for a in s.arrays()]
    g = Solver(grid, bed)
    g.stable_dt(s)
    for a, b in zip(s.arrays(), before):
        np.testing.assert_array_equal(a, b)
    g.step(s, 0.01)
    assert not s._host_valid  # the new state lives on the device only
    h = s.arrays()[0]
    assert s._host_valid and h.shape == before[0].shape


def test_invalid_setups_raise_like_the_reference_panics():
    grid = Grid(20, 10, 1.0, 1.0)
    bed = Bathymetry.flat(grid, 2.0)
    with pytest.raises(ValueError):
        Solver(grid, bed).with_shallow_cells([True] * 3)
    with pytest.raises(ValueError):  # maker on land
        Solver(grid, Bathymetry.flat(grid, -1.0)).with_wavemaker(
            WaveMaker(5.0, 0.1, 10.0, 0.0, 2.0, 1.0))
    with pytest.raises(ValueError):  # oblique wave that does not fit the period
        Solver(grid, bed).with_wavemaker(WaveMaker(5.0, 0.1, 10.0, 0.3, 2.0, 1.0)).with_periodic_y()
    with pytest.raises(ValueError):
        Solver(Grid(2, 10, 1.0, 1.0), Bathymetry.flat(Grid(2, 10, 1.0, 1.0), 1.0))


@pytest.mark.parametrize("pcg", [1, 3, 5])
@pytest.mark.parametrize("periodic", [False, True])
def test_breaking_fronts_and_eddy_viscosity(periodic, pcg):
    """Breaking::viscosity on (solver.rs:512-532, eddy.rs): a running hump as tall as the water
    is deep breaks at once; the fronts, their memory in the breaking array, the switch kept off
    behind them and the sub-stepped eddy viscosity must match the oracle bit for bit."""
    from wavesim_amd.config import Breaking, EddyViscosity

    grid = Grid(80, 24, 0.5, 0.5)
    bed = Bathymetry.flat(grid, 1.0)
    disp = Dispersion(breaking=Breaking(viscosity=EddyViscosity()), tolerance=1e-8)

    def configure(s):
        s = s.with_order(Order.Fifth).with_dispersion(disp)
        return s.with_periodic_y() if periodic else s

    s = bore_state(grid, 1.0)
    g, o, sg, so = lockstep(grid, bed, configure, s, 160,
                            f"eddy periodic={periodic} pcg={pcg}", options={"pcg_kernels": pcg})
    br = sg.arrays()[3]
    assert np.any(br > 0.0) and np.any(br < 0.0)  # fronts now, and lingering behind them
    ind = g.breaking_indicator(sg)
    assert np.array_equal(ind, o.breaking_indicator(so))
    assert np.all(ind[grid.inner(br).ravel() != 0.0] == 1.0)
