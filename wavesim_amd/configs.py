"""The synthetic workloads the measurements and the parity tests run.

Each builder returns a ``Workload``: grid, bed, the solver settings (applied to any solver with
wavecore's builder API, GPU or oracle) and the still-water start. Parameters the workload
descriptions leave open follow wavesim's planning conventions (crates/wavesim/src/lib.rs:437,
539-540, 644-645, 1706): wave-maker sigma = 2 dx, ramp over 4 periods, sponge strength
1.5 s^-1, maker amplitude = H/2.

Random beds draw from the 64-bit LCG of the reference's tests (dispersion.rs:642-651):
``x <- x * 6364136223846793005 + 1442695040888963407``, ``u = (x >> 11) / 2^53``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable

import numpy as np

from .config import Dispersion, Order, Sponge, WaveMaker, linear_wave
from .grid import GHOST, Bathymetry, Grid, State

SPONGE_STRENGTH = 1.5  # crates/wavesim/src/lib.rs:539
SIGMA_CELLS = 2.0  # crates/wavesim/src/lib.rs:540
RAMP_PERIODS = 4.0  # crates/wavesim/src/lib.rs:437


class Lcg:
    """Uniform draws in [0, 1) from the reference tests' LCG."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & self.MASK
        return lo + (hi - lo) * ((self.state >> 11) / float(1 << 53))


@dataclass
class Workload:
    name: str
    grid: Grid
    bed: Bathymetry
    configure: Callable  # solver -> solver
    steps: int
    notes: dict = field(default_factory=dict)

    def solver(self, factory):
        return self.configure(factory(self.grid, self.bed.copy()))

    def still_water(self) -> State:
        return State.lake_at_rest(self.grid, self.bed, 0.0)


def fourier_overlay(rng: Lcg, modes: int, lam: tuple[float, float], amp: tuple[float, float]):
    """``modes`` plane waves with wavelength, direction, amplitude and phase drawn in that order."""
    waves = []
    for _ in range(modes):
        wl = rng.uniform(*lam)
        theta = rng.uniform(0.0, 2.0 * math.pi)
        a = rng.uniform(*amp)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        k = 2.0 * math.pi / wl
        waves.append((a, k * math.cos(theta), k * math.sin(theta), ph))

    def f(x, y):
        out = np.zeros_like(x)
        for a, kx, ky, ph in waves:
            out += a * np.cos(kx * x + ky * y + ph)
        return out

    return f


def snapped_angle(angle: float, k: float, ly: float) -> tuple[float, int]:
    """The incidence nearest ``angle`` whose alongshore wavenumber fits a periodic width ``ly``
    (solver.rs:421-429 refuses any other), as the reference's periodic tests choose theirs
    (waves.rs:114-116)."""
    target = k * math.sin(angle) * ly / (2.0 * math.pi)
    best = None
    for n in (math.floor(target), math.ceil(target)):
        s = 2.0 * math.pi * n / (k * ly)
        if abs(s) <= 1.0:
            a = math.asin(s)
            if best is None or abs(a - angle) < abs(best[0] - angle):
                best = (a, n)
    return best


def c0_plane_beach() -> Workload:
    """256x64 cells at 2 m, plane beach 12 m deep at x = 0 rising to 2 m above water at
    x = 512 m, walls, wave-maker at 60 m (H 1 m, T 12 s, normal), 40 m west sponge, Manning
    0.02, MUSCL-MC, HLL, SSP-RK2, CFL 0.45, no dispersion, 2000 steps."""
    grid = Grid(256, 64, 2.0, 2.0)
    bed = Bathymetry.plane_beach(grid, 14.0 / 512.0, 12.0)
    maker = WaveMaker(60.0, 0.5, 12.0, 0.0, SIGMA_CELLS * grid.dx, RAMP_PERIODS)

    def configure(s):
        s = (s.with_order(Order.Second).with_wavemaker(maker)
             .with_sponge(Sponge(20, SPONGE_STRENGTH).edges(True, False, False, False))
             .with_manning(0.02))
        s.cfl = 0.45
        return s

    return Workload("plane_beach_256x64_fv2", grid, bed, configure, 2000)


def c1_rough_bed(nx: int = 1024, ny: int = 512, seed: int = 1, dx: float = 3.0,
                 steps: int = 4000, tolerance: float = 1e-8) -> Workload:
    """1024x512 cells at 3 m, periodic in y: a 1:40 plane from 20 m depth plus 16 random Fourier
    modes (wavelengths 30-300 m, amplitudes 0-0.4 m, seed 1), wave-maker H 1.5 m, T 14 s at
    10 degrees snapped to the periodic width, fifth-order blended reconstruction, SGN dispersion
    by PCG to 1e-8, breaking switch on, CFL 0.45, 4000 steps."""
    grid = Grid(nx, ny, dx, dx)
    rng = Lcg(seed)
    overlay = fourier_overlay(rng, 16, (30.0, 300.0), (0.0, 0.4))
    bed = Bathymetry.from_fn(grid, lambda x, y: -20.0 + x / 40.0 + overlay(x, y))
    period, height = 14.0, 1.5
    maker_x = 60.0
    # the depth the wave-maker will read (solver.rs:405-406), for the snapped angle
    i = min(max(round(maker_x / dx - 0.5), 0), nx - 1) + GHOST
    depth = -bed.b[grid.idx(i, GHOST + ny // 2)]
    k = linear_wave(2.0 * math.pi / period, depth, True).k
    angle, n_fit = snapped_angle(math.radians(10.0), k, ny * dx)
    maker = WaveMaker(maker_x, height / 2.0, period, angle, SIGMA_CELLS * dx, RAMP_PERIODS)
    # wavesim's plan keeps the sponge plus 4 sigma behind the maker (lib.rs:646-651)
    sponge_cells = max(4, int((maker_x - 4.0 * SIGMA_CELLS * dx) // dx))

    def configure(s):
        s = (s.with_order(Order.Fifth)
             .with_dispersion(Dispersion(tolerance=tolerance, max_iterations=200))
             .with_periodic_y()
             .with_wavemaker(maker)
             .with_sponge(Sponge(sponge_cells, SPONGE_STRENGTH).edges(True, False, False, False)))
        s.cfl = 0.45
        return s

    notes = {"angle_deg": math.degrees(angle), "alongshore_waves": n_fit,
             "maker_depth_m": depth, "sponge_cells": sponge_cells}
    return Workload(f"rough_bed_{nx}x{ny}_fifth_sgn", grid, bed, configure, steps, notes)
