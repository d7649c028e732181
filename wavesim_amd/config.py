"""Solver settings: the reference's configuration types, and their C-ABI layouts.

Mirrors the builder inputs of crates/wavecore/src/solver.rs:34-50 (``Order``),
dispersion.rs:43-106 (``Breaking``, ``Dispersion``, ``DispersionStats``) and forcing.rs:9-207
(``WaveMaker``, ``Sponge``, ``LinearWave``), with the defaults the reference gives them. The
``ctypes`` structures match ``include/wavecore_hip.h`` field for field.
"""

from __future__ import annotations

import ctypes
import enum
import math
from dataclasses import dataclass, field

from .grid import G


class Order(enum.IntEnum):
    """Spatial and temporal order (solver.rs:34-50)."""

    First = 0
    Second = 1
    Fifth = 2


@dataclass(frozen=True)
class EddyViscosity:
    """Breaking fronts followed from step to step and their eddy viscosity (eddy.rs:63-89),
    Kennedy et al.'s (2000) coefficients by default."""

    initial: float = 0.65
    settled: float = 0.15
    duration: float = 5.0
    mixing: float = 1.2
    linger: float = 5.0

    def span(self, h: float) -> float:
        """eddy.rs:104-106"""
        return self.duration * math.sqrt(h / G)


@dataclass(frozen=True)
class Breaking:
    """When to switch the dispersive terms off (dispersion.rs:43-74); `viscosity` follows the
    breaking fronts (None: the switch is read afresh at every step)."""

    eta_ratio: tuple[float, float] = (0.30, 0.55)
    slope: tuple[float, float] = (0.40, 0.65)
    viscosity: EddyViscosity | None = None


@dataclass(frozen=True)
class Dispersion:
    """SGN settings (dispersion.rs:76-97)."""

    breaking: Breaking = field(default_factory=Breaking)
    nonlinear: bool = True
    tolerance: float = 1e-4
    max_iterations: int = 200


@dataclass(frozen=True)
class DispersionStats:
    solves: int = 0
    iterations: int = 0
    unconverged: int = 0


@dataclass(frozen=True)
class WaveMaker:
    """Internal wave-maker (forcing.rs:9-33)."""

    x: float
    amplitude: float
    period: float
    angle: float
    sigma: float
    ramp_periods: float


@dataclass(frozen=True)
class Sponge:
    """Absorbing layer (forcing.rs:141-207)."""

    width: int
    strength: float
    level: float = 0.0
    west: bool = False
    east: bool = False
    south: bool = False
    north: bool = False

    def edges(self, west: bool, east: bool, south: bool, north: bool) -> "Sponge":
        return Sponge(self.width, self.strength, self.level, west, east, south, north)


@dataclass(frozen=True)
class LinearWave:
    k: float
    phase_speed: float
    group_speed: float


@dataclass(frozen=True)
class Limiting:
    """The cell that limits the time step (solver.rs:248-258); i, j count interior cells."""

    i: int
    j: int
    depth: float
    speed: float


def linear_wave(omega: float, h: float, dispersive: bool) -> LinearWave:
    """forcing.rs:53-75; raises ValueError where the reference panics."""
    if not dispersive:
        c = math.sqrt(G * h)
        return LinearWave(omega / c, c, c)
    nu = omega * omega * h / G
    if not nu < 3.0:
        raise ValueError(
            f"a wave of angular frequency {omega} is too short for the dispersive model in "
            f"{h} m of water"
        )
    kh = math.sqrt(nu / (1.0 - nu / 3.0))
    k = kh / h
    phase = omega / k
    return LinearWave(k, phase, phase / (1.0 + kh * kh / 3.0))


def manning_factor(h: float, speed: float, dt: float, n: float) -> float:
    """forcing.rs:294-297."""
    k = G * n * n / math.pow(h, 4.0 / 3.0)
    return 1.0 / (1.0 + dt * k * speed)


# ---- C-ABI layouts (include/wavecore_hip.h) ------------------------------------------------
class CWaveMaker(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("x", "amplitude", "period", "angle", "sigma",
                                               "ramp_periods")]


class CSponge(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint64),
        ("strength", ctypes.c_double),
        ("level", ctypes.c_double),
        ("west", ctypes.c_int32),
        ("east", ctypes.c_int32),
        ("south", ctypes.c_int32),
        ("north", ctypes.c_int32),
    ]


class CBreaking(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("eta_ratio_lo", "eta_ratio_hi", "slope_lo",
                                               "slope_hi")]


class CDispersion(ctypes.Structure):
    _fields_ = [
        ("breaking", CBreaking),
        ("nonlinear", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("max_iterations", ctypes.c_uint64),
    ]


class CEddy(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("initial", "settled", "duration", "mixing",
                                               "linger")]


def c_eddy(e: EddyViscosity) -> CEddy:
    return CEddy(e.initial, e.settled, e.duration, e.mixing, e.linger)


class CDispersionStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("solves", "iterations", "unconverged")]


class CLimiting(ctypes.Structure):
    _fields_ = [("i", ctypes.c_uint64), ("j", ctypes.c_uint64), ("depth", ctypes.c_double),
                ("speed", ctypes.c_double)]


class CLinearWave(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("k", "phase_speed", "group_speed")]


def c_wavemaker(m: WaveMaker) -> CWaveMaker:
    return CWaveMaker(m.x, m.amplitude, m.period, m.angle, m.sigma, m.ramp_periods)


def c_sponge(s: Sponge) -> CSponge:
    if s.width < 0:
        raise ValueError("sponge width must be non-negative")
    return CSponge(s.width, s.strength, s.level, int(s.west), int(s.east), int(s.south),
                   int(s.north))


def c_dispersion(d: Dispersion) -> CDispersion:
    b = d.breaking
    return CDispersion(
        CBreaking(b.eta_ratio[0], b.eta_ratio[1], b.slope[0], b.slope[1]),
        int(bool(d.nonlinear)),
        0,
        float(d.tolerance),
        int(d.max_iterations),
    )
