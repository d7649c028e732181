"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU FP64 oracle (libwavecore_oracle.so).

``OracleSolver`` has the same methods as the product's ``wavesim_amd.Solver`` (and the
reference's ``Solver``), so the tests can run one scenario through both and compare. Only
tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py import this module.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from wavesim_amd.config import (
    CDispersion,
    CDispersionStats,
    CLimiting,
    CLinearWave,
    CSponge,
    CWaveMaker,
    Dispersion,
    DispersionStats,
    Limiting,
    LinearWave,
    Order,
    Sponge,
    WaveMaker,
    c_dispersion,
    c_eddy,
    c_sponge,
    c_wavemaker,
)
from wavesim_amd.grid import Bathymetry, Grid, State

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwavecore_oracle.so")

DOT_REFERENCE = 0
DOT_CANONICAL = 1  # the GPU's 64x4-tile order (state sums; the 64x4-tile PCG variants)
DOT_CANONICAL16 = 2  # the GPU's default FP64 PCG: 64x16 tiles, running sums per cell
MATH_GLIBC = 0
MATH_PORTABLE = 1

_D = ctypes.c_double
_P = ctypes.POINTER(ctypes.c_double)
_U64 = ctypes.c_uint64
_I32 = ctypes.c_int32
_VP = ctypes.c_void_p


def build() -> None:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "wco_last_error": (ctypes.c_char_p, []),
            "wco_solver_new": (_I32, [_U64, _U64, _D, _D, _P, ctypes.POINTER(_VP)]),
            "wco_solver_free": (None, [_VP]),
            "wco_set_cfl": (_I32, [_VP, _D]),
            "wco_with_order": (_I32, [_VP, _I32]),
            "wco_with_manning": (_I32, [_VP, _D]),
            "wco_with_sponge": (_I32, [_VP, ctypes.POINTER(CSponge)]),
            "wco_with_wavemaker": (_I32, [_VP, ctypes.POINTER(CWaveMaker)]),
            "wco_with_dispersion": (_I32, [_VP, ctypes.POINTER(CDispersion)]),
            "wco_with_eddy_viscosity": (_I32, [_VP, ctypes.c_void_p]),
            "wco_eddy_follow": (_I32, [_U64, _U64, _D, _D, ctypes.c_void_p, _P, _P, _P, _P, _P,
                                       _D, _P]),
            "wco_eddy_diffuse": (_I32, [_U64, _U64, _D, _D, _P, _P, _P, _P, _D,
                                        ctypes.POINTER(_U64)]),
            "wco_with_periodic_y": (_I32, [_VP]),
            "wco_with_shallow_cells": (_I32, [_VP, ctypes.POINTER(ctypes.c_uint8), _U64]),
            "wco_bed": (_I32, [_VP, _P]),
            "wco_set_dot_order": (_I32, [_VP, _I32]),
            "wco_set_math": (_I32, [_VP, _I32]),
            "wco_set_threads": (_I32, [_VP, _I32]),
            "wco_set_warm_start": (_I32, [_VP, _I32, _I32]),
            "wco_ws_weights": (_I32, [_I32, _I32, _P]),
            "wco_ws_extrapolate": (_I32, [_I32, _I32, _I32, _U64, _P, _P]),
            "wco_state_set": (_I32, [_VP, _P, _P, _P, _P, _D]),
            "wco_state_get": (_I32, [_VP, _P, _P, _P, _P, _P]),
            "wco_stable_dt": (_I32, [_VP, _P]),
            "wco_step": (_I32, [_VP, _D]),
            "wco_limiting_cell": (_I32, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(CLimiting)]),
            "wco_breaking_indicator": (_I32, [_VP, _P]),
            "wco_dispersion_stats": (_I32, [_VP, ctypes.POINTER(_I32),
                                            ctypes.POINTER(CDispersionStats)]),
            "wco_state_volume": (_I32, [_VP, _P]),
            "wco_state_max_abs_momentum": (_I32, [_VP, _P]),
            "wco_state_min_depth": (_I32, [_VP, _P]),
            "wco_fill_boundaries": (_I32, [_VP]),
            "wco_relaxation_new": (_I32, [_VP, ctypes.POINTER(CSponge), _U64,
                                          ctypes.POINTER(_VP)]),
            "wco_relaxation_free": (None, [_VP]),
            "wco_relaxation_len": (_U64, [_VP]),
            "wco_relaxation_cells": (_I32, [_VP, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
            "wco_relaxation_apply": (_I32, [_VP, _D, _P]),
            "wco_stats_new": (_I32, [_VP, _D, _D, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(_VP)]),
            "wco_stats_free": (None, [_VP]),
            "wco_stats_record": (_I32, [_VP, _D]),
            "wco_stats_read": (_I32, [_VP, _P, ctypes.POINTER(_U64), _P, _P, _P, _P, _P, _P]),
            "wco_stats_harmonic": (_I32, [_VP, ctypes.POINTER(_I32), _P, _P]),
            "wco_stats_finish": (_I32, [_VP]),
            "wco_linear_wave": (_I32, [_D, _D, _I32, ctypes.POINTER(CLinearWave)]),
            "wco_manning_factor": (_D, [_D, _D, _D, _D]),
            "wco_face_values": (None, [_P, _D, _P, _P]),
            "wco_smoothstep": (_D, [_D, _D, _D]),
            "wco_hypot": (_D, [_D, _D, _I32]),
            "wco_switch_off": (_I32, [_P, _P, _P, _P, _U64, _P, _P]),
            "wco_pow43": (_D, [_D, _I32]),
            "wco_matvec": (_I32, [_U64, _U64, _D, _D, _I32, _P, _P, _P, _P, _P, _P, _P]),
            "wco_dot": (_D, [_U64, _U64, _I32, _P, _P, _P, _P]),
            "wco_dispersion_apply": (_I32, [_U64, _U64, _D, _D, _I32,
                                            ctypes.POINTER(CDispersion), _P, _P, _P, _P, _P, _P,
                                            _P, _I32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def _check(rc: int) -> None:
    if rc != 0:
        raise ValueError(lib().wco_last_error().decode())


class OracleSolver:
    """The oracle behind the reference's ``Solver`` API (solver.rs:260-429)."""

    def __init__(self, grid: Grid, bed: Bathymetry, *, dot_order: int = DOT_REFERENCE,
                 math: int = MATH_GLIBC, threads: int = 1):
        self.grid = grid
        b = np.ascontiguousarray(bed.b, dtype=np.float64)
        if b.size != grid.cells():
            raise ValueError("bed must be a padded array of grid.cells() values")
        h = _VP()
        _check(lib().wco_solver_new(grid.nx, grid.ny, grid.dx, grid.dy, ptr(b), ctypes.byref(h)))
        self._h = h
        self._cfl = 0.4
        self.order = Order.Second
        self.manning = 0.0
        self.periodic_y = False
        self._dispersion = None
        _check(lib().wco_set_dot_order(h, dot_order))
        _check(lib().wco_set_math(h, math))
        _check(lib().wco_set_threads(h, threads))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().wco_solver_free(self._h)
            self._h = None

    # -- builder -----------------------------------------------------------------------
    @property
    def cfl(self) -> float:
        return self._cfl

    @cfl.setter
    def cfl(self, v: float) -> None:
        self._cfl = float(v)
        _check(lib().wco_set_cfl(self._h, self._cfl))

    def with_order(self, order: Order) -> "OracleSolver":
        _check(lib().wco_with_order(self._h, int(order)))
        self.order = Order(order)
        return self

    def with_warm_start(self, order: int, window: int = 0) -> "OracleSolver":
        """The FP32 form's extrapolated warm start of the SGN solves (order 0: the reference's
        start from the previous solution; window: past steps weighed, at least the order). Set
        after the dispersion."""
        _check(lib().wco_set_warm_start(self._h, int(order), int(window)))
        return self

    def with_manning(self, n: float) -> "OracleSolver":
        _check(lib().wco_with_manning(self._h, float(n)))
        self.manning = float(n)
        return self

    def with_sponge(self, sponge: Sponge) -> "OracleSolver":
        c = c_sponge(sponge)
        _check(lib().wco_with_sponge(self._h, ctypes.byref(c)))
        return self

    def with_wavemaker(self, maker: WaveMaker) -> "OracleSolver":
        c = c_wavemaker(maker)
        _check(lib().wco_with_wavemaker(self._h, ctypes.byref(c)))
        return self

    def with_dispersion(self, d: Dispersion) -> "OracleSolver":
        c = c_dispersion(d)
        _check(lib().wco_with_dispersion(self._h, ctypes.byref(c)))
        if d.breaking.viscosity is not None:
            e = c_eddy(d.breaking.viscosity)
            _check(lib().wco_with_eddy_viscosity(self._h, ctypes.byref(e)))
        self._dispersion = d
        return self

    def with_periodic_y(self) -> "OracleSolver":
        _check(lib().wco_with_periodic_y(self._h))
        self.periodic_y = True
        return self

    def with_shallow_cells(self, cells) -> "OracleSolver":
        a = np.ascontiguousarray(np.asarray(cells, dtype=bool).astype(np.uint8).ravel())
        _check(lib().wco_with_shallow_cells(
            self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), a.size))
        return self

    def dispersion(self):
        return self._dispersion

    @property
    def bed(self) -> Bathymetry:
        out = np.empty(self.grid.cells())
        _check(lib().wco_bed(self._h, ptr(out)))
        return Bathymetry(out)

    # -- state traffic -------------------------------------------------------------------
    def _upload(self, s: State) -> None:
        h, hu, hv, br = s.arrays()
        _check(lib().wco_state_set(self._h, ptr(np.ascontiguousarray(h)),
                                   ptr(np.ascontiguousarray(hu)), ptr(np.ascontiguousarray(hv)),
                                   ptr(np.ascontiguousarray(br)), s.time))

    def _download_into(self, s: State) -> None:
        n = self.grid.cells()
        out = [np.empty(n) for _ in range(4)]
        t = ctypes.c_double()
        _check(lib().wco_state_get(self._h, *(ptr(a) for a in out), ctypes.byref(t)))
        s._h, s._hu, s._hv, s._breaking = out
        s._host_valid = True
        s._dev_valid = False

    # -- the API ---------------------------------------------------------------------------
    def stable_dt(self, s: State) -> float:
        self._upload(s)
        out = ctypes.c_double()
        _check(lib().wco_stable_dt(self._h, ctypes.byref(out)))
        return out.value

    def step(self, s: State, dt: float) -> None:
        self._upload(s)
        _check(lib().wco_step(self._h, float(dt)))
        t = s.time
        self._download_into(s)
        s.time = t + float(dt)

    def limiting_cell(self, s: State):
        self._upload(s)
        found = ctypes.c_int32()
        out = CLimiting()
        _check(lib().wco_limiting_cell(self._h, ctypes.byref(found), ctypes.byref(out)))
        if not found.value:
            return None
        return Limiting(int(out.i), int(out.j), out.depth, out.speed)

    def breaking_indicator(self, s: State) -> np.ndarray:
        self._upload(s)
        out = np.empty(self.grid.nx * self.grid.ny)
        _check(lib().wco_breaking_indicator(self._h, ptr(out)))
        return out

    def dispersion_stats(self):
        has = ctypes.c_int32()
        out = CDispersionStats()
        _check(lib().wco_dispersion_stats(self._h, ctypes.byref(has), ctypes.byref(out)))
        if not has.value:
            return None
        return DispersionStats(int(out.solves), int(out.iterations), int(out.unconverged))

    def volume(self, s: State) -> float:
        """State::volume in this oracle's summation order."""
        self._upload(s)
        out = ctypes.c_double()
        _check(lib().wco_state_volume(self._h, ctypes.byref(out)))
        return out.value

    def step_n(self, s: State, steps: int) -> None:
        """``steps`` iterations of ``dt = stable_dt; step(dt)`` without host round trips."""
        self._upload(s)
        dt = ctypes.c_double()
        t = s.time
        for _ in range(steps):
            _check(lib().wco_stable_dt(self._h, ctypes.byref(dt)))
            if not (np.isfinite(dt.value) and dt.value > 0.0):
                raise FloatingPointError(f"no valid time step: {dt.value}")
            _check(lib().wco_step(self._h, dt.value))
            t = t + dt.value
        self._download_into(s)
        s.time = t


def relaxation_targets(target, n: int) -> np.ndarray:
    """The (eta, u, v) of each of n zone cells, interleaved, from a callable m -> (eta, u, v)
    (the reference's closure) or an array of shape (n, 3)."""
    if callable(target):
        out = np.empty((n, 3))
        for m in range(n):
            out[m] = target(m)
    else:
        out = np.array(target, dtype=np.float64).reshape(n, 3)
    return np.ascontiguousarray(out.reshape(-1))


class OracleRelaxation:
    """Relaxation (forcing.rs:227-284) on the oracle: ``new``, ``cells``, ``apply``."""

    def __init__(self, solver: OracleSolver, zones):
        self.solver = solver
        arr = (CSponge * max(len(zones), 1))(*[c_sponge(z) for z in zones])
        h = _VP()
        _check(lib().wco_relaxation_new(solver._h, arr, len(zones), ctypes.byref(h)))
        self._r = h

    def __del__(self):
        if getattr(self, "_r", None):
            lib().wco_relaxation_free(self._r)
            self._r = None

    def __len__(self) -> int:
        return int(lib().wco_relaxation_len(self._r))

    def cells(self) -> list[tuple[int, int]]:
        n = len(self)
        i = (_U64 * max(n, 1))()
        j = (_U64 * max(n, 1))()
        _check(lib().wco_relaxation_cells(self._r, i, j))
        return [(int(i[m]), int(j[m])) for m in range(n)]

    def apply(self, s: State, dt: float, target) -> None:
        t = relaxation_targets(target, len(self))
        self.solver._upload(s)
        _check(lib().wco_relaxation_apply(self._r, float(dt), ptr(t)))
        time = s.time
        self.solver._download_into(s)
        s.time = time


class OracleStats:
    """wavesim's Stats and Stage::record (crates/wavesim/src/lib.rs:899-932, 1276-1300) on the
    oracle solver's state (record after the oracle has stepped the State it holds)."""

    def __init__(self, solver: OracleSolver, from_: float, period: float, elevation=None):
        self.solver = solver
        self.n = solver.grid.nx * solver.grid.ny
        el = None
        if elevation is not None:
            self._el = np.ascontiguousarray(elevation, dtype=np.float32)
            el = self._el.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        h = _VP()
        _check(lib().wco_stats_new(solver._h, float(from_), float(period), el, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().wco_stats_free(self._h)
            self._h = None

    def record(self, s: State, dt: float) -> None:
        self.solver._upload(s)
        _check(lib().wco_stats_record(self._h, float(dt)))

    def finish(self) -> None:
        """Stage::finish (lib.rs:1451-1453), after the last record."""
        _check(lib().wco_stats_finish(self._h))

    def read(self) -> dict:
        keys = ("breaking_seconds", "eta_max", "eta_min", "wave_mean", "wave_first", "wave_last")
        out = {k: np.empty(self.n) for k in keys}
        sec = ctypes.c_double()
        waves = _U64()
        _check(lib().wco_stats_read(self._h, ctypes.byref(sec), ctypes.byref(waves),
                                    *(ptr(out[k]) for k in keys)))
        out["seconds"] = sec.value
        out["waves"] = int(waves.value)
        return out

    def harmonic(self):
        ok = _I32()
        a, b = np.empty(self.n), np.empty(self.n)
        _check(lib().wco_stats_harmonic(self._h, ctypes.byref(ok), ptr(a), ptr(b)))
        return (a, b) if ok.value else None


def linear_wave(omega: float, h: float, dispersive: bool) -> LinearWave:
    out = CLinearWave()
    _check(lib().wco_linear_wave(omega, h, int(dispersive), ctypes.byref(out)))
    return LinearWave(out.k, out.phase_speed, out.group_speed)


def manning_factor(h: float, speed: float, dt: float, n: float) -> float:
    return lib().wco_manning_factor(h, speed, dt, n)


def face_values(v, blend: float) -> tuple[float, float]:
    a = np.ascontiguousarray(v, dtype=np.float64)
    lo, hi = ctypes.c_double(), ctypes.c_double()
    lib().wco_face_values(ptr(a), blend, ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def smoothstep(lo: float, hi: float, x: float) -> float:
    return lib().wco_smoothstep(lo, hi, x)


def hypot(a: float, b: float, portable: bool) -> float:
    return lib().wco_hypot(a, b, int(portable))


def pow43(h: float, portable: bool) -> float:
    return lib().wco_pow43(h, int(portable))


def switch_off(thresholds, ratio, ex, ey):
    """The GPU switch's shortcut (hypot skipped where its bits cannot matter) and the full
    expression, cell by cell: (shortcut, full)."""
    th = np.ascontiguousarray(thresholds, dtype=np.float64)
    ratio, ex, ey = (np.ascontiguousarray(a, dtype=np.float64) for a in (ratio, ex, ey))
    fast, full = np.empty_like(ratio), np.empty_like(ratio)
    lib().wco_switch_off(ptr(th), ptr(ratio), ptr(ex), ptr(ey), ratio.size, ptr(fast),
                         ptr(full))
    return fast, full


def matvec(grid: Grid, periodic: bool, heff, mask, c, px, py):
    n = grid.cells()
    ax, ay = np.zeros(n), np.zeros(n)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (heff, mask, c, px, py)]
    _check(lib().wco_matvec(grid.nx, grid.ny, grid.dx, grid.dy, int(periodic),
                            *(ptr(a) for a in args), ptr(ax), ptr(ay)))
    return ax, ay


def dot(grid: Grid, ax, ay, bx, by, order: int = DOT_REFERENCE) -> float:
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (ax, ay, bx, by)]
    return lib().wco_dot(grid.nx, grid.ny, order, *(ptr(a) for a in args))


def dispersion_apply(grid: Grid, periodic: bool, d: Dispersion, h, hu, hv, phi, lh, lhu, lhv,
                     dot_order: int = DOT_REFERENCE):
    """Dispersion::apply with fresh scratch; returns the new (lhu, lhv)."""
    c = c_dispersion(d)
    ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (h, hu, hv, phi, lh)]
    lhu = np.array(lhu, dtype=np.float64)
    lhv = np.array(lhv, dtype=np.float64)
    _check(lib().wco_dispersion_apply(grid.nx, grid.ny, grid.dx, grid.dy, int(periodic),
                                      ctypes.byref(c), *(ptr(a) for a in ins), ptr(lhu),
                                      ptr(lhv), dot_order))
    return lhu, lhv


class OracleSampler:
    """wavesim's Sampler (crates/wavesim/src/lib.rs:1548-1616) and the nested run's time
    interpolation of its targets (lib.rs:1877-1890), in plain Python floats (IEEE doubles, the
    reference's operation order). ``cells`` are the zone's padded (i, j)."""

    def __init__(self, cells, fine_x0, fine_dx, fine_dy, cnx, cny, cdx, cdy, crop=0):
        import math

        g = 3  # GHOST
        w = cnx + 2 * g
        self.corners = []
        for i, j in cells:
            x = float(fine_x0 + i - g) * fine_dx + 0.5 * fine_dx
            y = float(j - g) * fine_dy + 0.5 * fine_dy
            cx = x / cdx - float(crop) - 0.5
            cy = y / cdy - 0.5
            ix, iy = math.floor(cx), math.floor(cy)
            wx, wy = cx - ix, cy - iy

            def at(di, dj, ix=ix, iy=iy):
                pi = min(max(int(ix + di) + g, g - 1), cnx + g)
                pj = min(max(int(iy + dj) + g, g - 1), cny + g)
                return pj * w + pi

            self.corners.append([(at(0.0, 0.0), (1.0 - wx) * (1.0 - wy)),
                                 (at(1.0, 0.0), wx * (1.0 - wy)),
                                 (at(0.0, 1.0), (1.0 - wx) * wy),
                                 (at(1.0, 1.0), wx * wy)])

    def sample(self, h, hu, hv, b):
        """(eta, u, v) per zone cell from padded arrays whose ghosts are filled (walls)."""
        out = []
        for corners in self.corners:
            eta = u = v = weight = 0.0
            for k, wk in corners:
                if h[k] > 1e-3 and wk > 0.0:
                    eta += wk * (h[k] + b[k])
                    u += wk * hu[k] / h[k]
                    v += wk * hv[k] / h[k]
                    weight += wk
            out.append((eta / weight, u / weight, v / weight) if weight > 0.0
                       else (float("-inf"), 0.0, 0.0))
        return np.array(out)

    @staticmethod
    def lerp(frm, to, w):
        out = np.empty_like(frm)
        for m in range(len(frm)):
            a, c = frm[m], to[m]
            if np.isfinite(a[0]) and np.isfinite(c[0]):
                out[m] = (a[0] + w * (c[0] - a[0]), a[1] + w * (c[1] - a[1]),
                          a[2] + w * (c[2] - a[2]))
            else:
                out[m] = a if w < 0.5 else c
        return out


def eddy_follow(grid, e, h, hu, hv, breaking, phi, dt):
    """EddyViscosity::follow (eddy.rs:163-221) on padded arrays: (new breaking, nu)."""
    h, hu, hv, phi = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (h, hu, hv, phi))
    br = np.ascontiguousarray(breaking, dtype=np.float64).copy()
    nu = np.zeros_like(h)
    ce = c_eddy(e)
    _check(lib().wco_eddy_follow(grid.nx, grid.ny, grid.dx, grid.dy, ctypes.byref(ce), ptr(h),
                                 ptr(hu), ptr(hv), ptr(br), ptr(phi), float(dt), ptr(nu)))
    return br, nu


def eddy_diffuse(grid, h, hu, hv, nu, dt):
    """diffuse (eddy.rs:225-302) on padded arrays, walls refilled between sub-steps:
    (hu, hv, sub-steps)."""
    h, hu, hv, nu = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (h, hu, hv, nu))
    n = _U64()
    _check(lib().wco_eddy_diffuse(grid.nx, grid.ny, grid.dx, grid.dy, ptr(h), ptr(hu), ptr(hv),
                                  ptr(nu), float(dt), ctypes.byref(n)))
    return hu, hv, int(n.value)


def ws_extrapolate(order: int, period: int, past, window: int = 0) -> float:
    """The warm start of solve n for one cell from past[i] = x_{n-1-i}, as the oracle's solver
    (and the FP32 form) computes it."""
    a = np.ascontiguousarray(past, dtype=np.float64)
    out = ctypes.c_double()
    _check(lib().wco_ws_extrapolate(int(order), int(window), int(period), a.size,
                                    a.ctypes.data_as(_P), ctypes.byref(out)))
    return out.value


def ws_weights(order: int, window: int = 0) -> np.ndarray:
    """The oracle's warm-start weights c_1..c_K, K = max(order, window)."""
    out = np.zeros(max(order, window))
    _check(lib().wco_ws_weights(int(order), int(window), out.ctypes.data_as(_P)))
    return out
