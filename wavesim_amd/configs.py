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


def _sponge_cells(period: float, depth: float, dx: float, nx: int) -> int:
    """wavesim's offshore sponge: 1.5 wavelengths, clamped to [4, nx/3] (lib.rs:644-645)."""
    wl = 2.0 * math.pi / linear_wave(2.0 * math.pi / period, depth, True).k
    return int(min(max(math.ceil(1.5 * wl / dx), 4), max(nx // 3, 4)))


def gaussian_lobes(rng: Lcg, n: int, xr, yr, height, width):
    lobes = [(rng.uniform(*xr), rng.uniform(*yr), rng.uniform(*height), rng.uniform(*width))
             for _ in range(n)]

    def f(x, y):
        out = np.zeros_like(x)
        for x0, y0, hgt, wd in lobes:
            out += hgt * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2.0 * wd * wd))
        return out

    return f


def c3_reef_instance(seed: int, dx: float = 3.0, size: tuple[int, int] | None = None) -> Workload:
    """One instance of the batched spot-sheet sweep: a size drawn from {256x128, 512x256,
    768x384, 1024x512, 1536x768}, a 1:30 plane (levelling off at 40 m depth, so that the
    wave-maker of the larger grids stays within the dispersive model's range) plus 1-4 Gaussian
    reef lobes (4-10 m high, 40-200 m wide) plus a 3 m channel, swell angle -20..20 degrees,
    tide -0.5..0.5 m (the bed shifted down by the tide, as waveio's Bed::bathymetry does,
    crates/waveio/src/bed.rs:239-246), H 1-3 m, T 10-18 s; fifth order + SGN, 2000 steps.
    `size` replaces the drawn size (the rest of the draw is unchanged)."""
    rng = Lcg(seed)
    sizes = [(256, 128), (512, 256), (768, 384), (1024, 512), (1536, 768)]
    nx, ny = sizes[min(int(rng.uniform() * len(sizes)), len(sizes) - 1)]
    if size is not None:
        nx, ny = size
    grid = Grid(nx, ny, dx, dx)
    lx, ly = nx * dx, ny * dx
    angle = math.radians(rng.uniform(-20.0, 20.0))
    tide = rng.uniform(-0.5, 0.5)
    height = rng.uniform(1.0, 3.0)
    period = rng.uniform(10.0, 18.0)
    n_lobes = 1 + min(int(rng.uniform() * 4), 3)
    # the plane meets still water 85% of the way across; the offshore depth follows
    shore = 0.85 * lx
    lobes = gaussian_lobes(rng, n_lobes, (0.45 * lx, 0.8 * lx), (0.0, ly), (4.0, 10.0),
                           (40.0, 200.0))
    ch_y = rng.uniform(0.2 * ly, 0.8 * ly)

    def bed_fn(x, y):
        plane = np.maximum((x - shore) / 30.0, -40.0)
        channel = -3.0 * np.exp(-((y - ch_y) / 30.0) ** 2) * (x > 0.4 * lx)
        return np.minimum(plane + lobes(x, y), plane + 2.0) + channel - tide

    bed = Bathymetry.from_fn(grid, bed_fn)
    depth0 = float(-np.max(grid.inner(bed.b)[:, 0]))
    sponge = _sponge_cells(period, max(depth0, 1.0), dx, nx)
    maker_x = (sponge + 4.0 * SIGMA_CELLS) * dx
    maker = WaveMaker(maker_x, height / 2.0, period, angle, SIGMA_CELLS * dx, RAMP_PERIODS)

    # the side sponges at most an eighth of the sheet's width each: 1.5 wavelengths on both
    # sides of a 384 m sheet would damp all of it, and no wave would reach the reef
    side = min(sponge, max(ny // 8, 4))

    def configure(s):
        s = (s.with_order(Order.Fifth)
             .with_dispersion(Dispersion(tolerance=1e-8, max_iterations=200))
             .with_wavemaker(maker)
             .with_sponge(Sponge(sponge, SPONGE_STRENGTH).edges(True, False, False, False))
             .with_sponge(Sponge(side, SPONGE_STRENGTH).edges(False, False, True, True)))
        s.cfl = 0.45
        return s

    notes = {"seed": seed, "nx": nx, "ny": ny, "angle_deg": math.degrees(angle), "tide_m": tide,
             "H_m": height, "T_s": period, "lobes": n_lobes, "sponge_cells": sponge,
             "side_sponge_cells": side}
    return Workload(f"reef_sheet_{seed}_{nx}x{ny}", grid, bed, configure, 2000, notes)


def c4_large_reef(nx: int = 4096, ny: int = 2048, dx: float = 3.0, seed: int = 4000,
                  steps: int = 6000) -> Workload:
    """4096x2048 cells at 3 m: a shelf 40 m deep shoaling at 1:50 onto a reef flat 1 m deep,
    8 Gaussian reef lobes, the rough-bed Fourier overlay (16 modes, 30-300 m, 0-0.4 m), a dry
    beach strip; wave-maker H 3 m, T 16 s at 15 degrees, sponges on the west, south and north,
    Manning 0.025, fifth order + SGN by PCG to 1e-8, breaking switch on, 6000 steps.

    Layout: the west sponge and two wavelengths of the 40 m shelf, then the 1:50 slope (1950 m)
    and the reef flat to the beach strip at 95% of the domain. The wave-maker sits on the slope
    at the 15 m contour, so that the waves reach the reef edge and break there within the
    run's 6000 steps (about 190 s of simulated time); from the 40 m shelf they would need more
    than 200 s to get there."""
    grid = Grid(nx, ny, dx, dx)
    lx, ly = nx * dx, ny * dx
    rng = Lcg(seed)
    overlay = fourier_overlay(rng, 16, (30.0, 300.0), (0.0, 0.4))
    period, height = 16.0, 3.0
    sponge = _sponge_cells(period, 40.0, dx, nx)
    wavelength = 2.0 * math.pi / linear_wave(2.0 * math.pi / period, 40.0, True).k
    x_edge = sponge * dx + 2.0 * wavelength  # shelf edge
    x_flat = x_edge + 39.0 * 50.0  # 40 m -> 1 m at 1:50
    x_beach = 0.95 * lx  # dry beach strip beyond
    lobes = gaussian_lobes(rng, 8, (x_edge, x_beach), (0.0, ly), (4.0, 10.0), (40.0, 200.0))

    def bed_fn(x, y):
        base = np.where(x < x_edge, -40.0,
                        np.where(x < x_flat, -40.0 + (x - x_edge) / 50.0, -1.0))
        base = np.where(x > x_beach, -1.0 + (x - x_beach) / 20.0, base)
        return np.minimum(base + lobes(x, y), 3.0) + overlay(x, y)

    bed = Bathymetry.from_fn(grid, bed_fn)
    maker_x = x_edge + (40.0 - 15.0) * 50.0  # the 15 m contour of the slope
    maker = WaveMaker(maker_x, height / 2.0, period, math.radians(15.0), SIGMA_CELLS * dx,
                      RAMP_PERIODS)

    def configure(s):
        s = (s.with_order(Order.Fifth)
             .with_dispersion(Dispersion(tolerance=1e-8, max_iterations=200))
             .with_wavemaker(maker)
             .with_sponge(Sponge(sponge, SPONGE_STRENGTH).edges(True, False, True, True))
             .with_manning(0.025))
        s.cfl = 0.45
        return s

    notes = {"seed": seed, "sponge_cells": sponge, "maker_x_m": maker_x}
    return Workload(f"large_reef_{nx}x{ny}_fifth_sgn", grid, bed, configure, steps, notes)


def c4_crop(i0: int = 700, j0: int = 960, nx: int = 256, ny: int = 96, sponge: int = 8,
            steps: int = 200) -> Workload:
    """A tile of the large reef: interior columns [i0, i0+nx) and rows [j0, j0+ny) of its bed
    (by default the slope around the wave-maker, 15 m contour, up to the 4 m contour where the
    waves break), with its physics: fifth order, SGN to 1e-8, Manning 0.025, the same wave-maker
    at its place in the tile (the angle snapped to the tile's width, as wavesim snaps it), and
    sponges `sponge` cells wide on the tile's west, south and north edges."""
    full = c4_large_reef()
    g = full.grid
    inner = full.grid.inner(full.bed.b)[j0:j0 + ny, i0:i0 + nx]
    grid = Grid(nx, ny, g.dx, g.dy)
    bed = Bathymetry.from_interior(grid, inner)
    maker_x = full.notes["maker_x_m"] - i0 * g.dx
    if not 0.0 < maker_x < nx * g.dx:
        raise ValueError("the crop must hold the wave-maker")
    i = round(maker_x / g.dx - 0.5) + GHOST
    depth = -bed.b[grid.idx(i, GHOST + ny // 2)]
    period = 16.0
    k = linear_wave(2.0 * math.pi / period, depth, True).k
    # periodic in y is not set, so any angle is allowed; keep 15 degrees
    maker = WaveMaker(maker_x, 1.5, period, math.radians(15.0), SIGMA_CELLS * g.dx,
                      RAMP_PERIODS)

    def configure(s):
        s = (s.with_order(Order.Fifth)
             .with_dispersion(Dispersion(tolerance=1e-8, max_iterations=200))
             .with_wavemaker(maker)
             .with_sponge(Sponge(sponge, SPONGE_STRENGTH).edges(True, False, True, True))
             .with_manning(0.025))
        s.cfl = 0.45
        return s

    notes = {"i0": i0, "j0": j0, "maker_depth_m": depth, "k": k}
    return Workload(f"large_reef_crop_{nx}x{ny}", grid, bed, configure, steps, notes)
