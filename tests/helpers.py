"""Run loops shared by the tests, written like the reference's own test drivers.

``run`` follows crates/wavecore/tests/well_balanced.rs:7-13, ``run_to`` tests/dispersion.rs:21-27
and ``run_and_analyse`` / ``Harmonic`` tests/common/mod.rs:5-68. They take any solver with the
reference's ``stable_dt`` / ``step`` API: the CPU oracle or the GPU path.
"""

from __future__ import annotations

import math

import numpy as np

from wavesim_amd.grid import GHOST, Grid, State


def run(solver, s, steps: int) -> None:
    for _ in range(steps):
        dt = solver.stable_dt(s)
        assert math.isfinite(dt) and dt > 0.0, "no valid time step"
        solver.step(s, dt)


def run_to(solver, s, t_end: float) -> None:
    while s.time < t_end:
        dt = solver.stable_dt(s)
        assert math.isfinite(dt), "no water in the domain"
        solver.step(s, min(dt, t_end - s.time))


class Harmonic:
    """Complex amplitude at angular frequency ``omega`` per probe (tests/common/mod.rs:5-40)."""

    def __init__(self, omega: float, window: tuple[float, float], probes: int):
        self.omega = omega
        self.window = window
        self.sums = [[0.0, 0.0] for _ in range(probes)]

    def record(self, t: float, dt: float, eta) -> None:
        if t <= self.window[0] or t > self.window[1]:
            return
        c, s = math.cos(self.omega * t), math.sin(self.omega * t)
        for acc, e in zip(self.sums, eta):
            acc[0] += e * c * dt
            acc[1] -= e * s * dt

    def result(self):
        scale = 2.0 / (self.window[1] - self.window[0])
        return [(scale * math.hypot(re, im), math.atan2(im, re)) for re, im in self.sums]


def run_and_analyse(solver, s, t_end, omega, window, probes, grid: Grid, bed_b):
    """tests/common/mod.rs:43-68; ``probes`` are padded (i, j)."""
    harmonic = Harmonic(omega, window, len(probes))
    ks = [grid.idx(i, j) for i, j in probes]
    while s.time < t_end:
        dt = solver.stable_dt(s)
        assert math.isfinite(dt), "no water in the domain: nothing can propagate"
        dt = min(dt, t_end - s.time)
        solver.step(s, dt)
        if s.time > window[0]:
            h = s.arrays()[0]
            harmonic.record(s.time, dt, [h[k] + bed_b[k] for k in ks])
    return harmonic.result()


def col(grid: Grid, x: float) -> int:
    return GHOST + int(x / grid.dx)


class Lcg:
    """Deterministic pseudo-random numbers in [-1, 1) (dispersion.rs:642-651)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed

    def next(self) -> float:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & self.MASK
        return ((self.state >> 11) / float(1 << 53)) * 2.0 - 1.0


def max_rel_diff(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max(|b|), NaN-aware: a NaN on one side only is an infinite difference."""
    a = np.asarray(a)
    b = np.asarray(b)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    if np.any(nan_a != nan_b):
        return math.inf
    ok = ~nan_a
    scale = float(np.max(np.abs(b[ok]), initial=0.0))
    diff = float(np.max(np.abs(a[ok] - b[ok]), initial=0.0))
    if scale == 0.0:
        return diff
    return diff / scale


def bore_state(grid: Grid, depth: float, amp=1.0, x0=12.0, y0=4.0, rx=2.5, ry=5.0):
    """A hump as tall as the water is deep on a flat bed, running towards +x with the
    shallow-water Riemann invariant u = 2 (sqrt(g h) - sqrt(g depth)): its front steepens and
    breaks within the first steps, so fronts (breaking > 0) and lingering cells (breaking < 0)
    are both present for most of 160 steps (an oracle run of this state shows them at step
    160, for walls and for the periodic seam)."""
    s = State.zeros(grid)
    x, y = grid.centres()
    X, Y = np.meshgrid(x[GHOST:GHOST + grid.nx], y[GHOST:GHOST + grid.ny])
    h = depth + amp * np.exp(-(X - x0) ** 2 / (2 * rx * rx) - (Y - y0) ** 2 / (2 * ry * ry))
    grid.inner(s.h)[:, :] = h
    grid.inner(s.hu)[:, :] = h * 2.0 * (np.sqrt(9.81 * h) - np.sqrt(9.81 * depth))
    s.fill_walls(grid)
    return s
