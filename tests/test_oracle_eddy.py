"""Breaking fronts and their eddy viscosity (crates/wavecore/src/eddy.rs) on the oracle: the
reference's own unit tests (eddy.rs:309-442) through the oracle's follow and diffuse, plus the
closed forms they leave open."""

from __future__ import annotations

import math

import numpy as np

from oracle import oracle as orc
from wavesim_amd.config import EddyViscosity
from wavesim_amd.grid import GHOST, G, Grid, State


def channel(grid, breaking, momentum):
    """Still water 1 m deep, hu set by momentum(i) along x (eddy.rs:322-333)."""
    s = State.zeros(grid)
    for i, j in grid.interior():
        k = grid.idx(i, j)
        s.h[k] = 1.0
        s.hu[k] = momentum(i)
    s.breaking[:] = breaking
    s.fill_walls(grid)
    return s


def threshold(e, age, h):
    span = e.span(h)
    return e.settled if age >= span else e.initial + (e.settled - e.initial) * age / span


def test_the_threshold_falls_from_onset_to_settled_over_the_duration():
    e = EddyViscosity()
    span = 5.0 * math.sqrt(1.0 / G)
    assert threshold(e, 0.0, 1.0) == 0.65
    assert abs(threshold(e, 0.5 * span, 1.0) - 0.40) < 1e-12
    assert threshold(e, 2.0 * span, 1.0) == 0.15


def test_diffusion_conserves_momentum_and_smooths_a_jump():
    grid = Grid(40, 3, 0.1, 0.1)
    s = channel(grid, np.zeros(grid.cells()), lambda i: 0.5 if i < GHOST + 20 else 0.0)
    nu = np.zeros(grid.cells())
    for i, j in grid.interior():
        if GHOST + 8 <= i < GHOST + 32:
            nu[grid.idx(i, j)] = 0.05
    inner = [grid.idx(i, j) for i, j in grid.interior()]
    before = sum(s.hu[k] for k in inner)
    hu, hv, steps = orc.eddy_diffuse(grid, s.h, s.hu, s.hv, nu, 0.5)
    assert steps > 1
    assert abs(sum(hu[k] for k in inner) - before) < 1e-9 * before
    j = GHOST + 1
    a, b = hu[grid.idx(GHOST + 19, j)], hu[grid.idx(GHOST + 20, j)]
    assert a < 0.5 and b > 0.0 and a > b
    assert all(-1e-12 <= hu[k] <= 0.5 + 1e-12 for k in inner)
    # the sub-step count is the reference's stability bound (eddy.rs:247-251)
    idx2 = idy2 = 1.0 / 0.01
    limit = 0.5 / (0.05 * max(2.0 * idx2 + idy2, idx2 + 2.0 * idy2))
    assert steps == max(math.ceil(0.5 / limit), 1)


def test_a_uniform_viscosity_diffuses_a_sine_at_the_heat_equation_rate():
    # d(hu)/dt = nu d2(hu)/dx2 along x: a mode sin(k x) decays by the discrete factor per
    # sub-step, (1 - sub nu 4 sin^2(k dx / 2) / dx^2)
    grid = Grid(32, 3, 1.0, 1.0)
    k = 2 * math.pi / 32.0
    s = channel(grid, np.zeros(grid.cells()),
                lambda i: 0.2 * math.cos(k * ((i - GHOST) + 0.5)))
    nu = np.zeros(grid.cells())
    for i, j in grid.interior():
        nu[grid.idx(i, j)] = 0.1
    hu, _, steps = orc.eddy_diffuse(grid, s.h, s.hu, s.hv, nu, 1.0)
    sub = 1.0 / steps
    f = (1.0 - sub * 0.1 * 4.0 * math.sin(k / 2.0) ** 2) ** steps
    j = GHOST + 1
    for i in range(GHOST + 4, GHOST + 28):
        want = f * 0.2 * math.cos(k * ((i - GHOST) + 0.5))
        assert abs(hu[grid.idx(i, j)] - want) < 1e-12


def test_a_front_starts_where_the_switch_says_moves_on_with_its_age_and_lingers_behind():
    grid = Grid(20, 3, 0.1, 0.1)
    e = EddyViscosity()
    dt, c = 0.01, math.sqrt(G)

    def cell(i):
        return grid.idx(GHOST + i, GHOST + 1)

    def towards(at, rate):
        a = rate * c * 0.1
        return lambda i: a if i < GHOST + at else (0.0 if i == GHOST + at else -a)

    smooth = np.ones(grid.cells())
    none = np.zeros(grid.cells())

    def follow(s, phi):
        return orc.eddy_follow(grid, e, s.h, s.hu, s.hv, s.breaking, phi, dt)

    br, _ = follow(channel(grid, none, towards(5, 0.8)), smooth)
    assert np.all(br == 0.0)
    broken = smooth.copy()
    for i in range(3, 8):
        broken[cell(i)] = 0.0
    br, nu = follow(channel(grid, none, towards(5, 0.8)), broken)
    assert br[cell(5)] == dt and br[cell(4)] == dt and br[cell(3)] == 0.0
    assert nu[cell(5)] > 0.0 and nu[cell(4)] == 0.0
    old = 2.0 * e.span(1.0)
    bore = none.copy()
    for i in range(4, 7):
        bore[cell(i)] = old
    br, _ = follow(channel(grid, bore, towards(6, 0.7)), smooth)
    assert br[cell(6)] == old + dt and br[cell(5)] == old + dt
    linger = e.linger * math.sqrt(1.0 / G)
    assert br[cell(4)] == -linger
    br2, nu2 = follow(channel(grid, br, lambda i: 0.0), smooth)
    assert br2[cell(6)] == -linger
    assert abs(br2[cell(4)] + linger - dt) < 1e-15
    assert np.all(nu2 == 0.0)
