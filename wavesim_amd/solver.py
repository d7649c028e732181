"""``Solver``: wavecore's Solver (crates/wavecore/src/solver.rs:260-507) on the GPU.

Same builder methods, same ``stable_dt`` / ``step`` / ``limiting_cell`` / ``breaking_indicator``
/ ``dispersion_stats``, same errors (``ValueError`` where the reference panics). Every call goes
through the C ABI of ``libwavecore_hip.so`` (include/wavecore_hip.h); the state stays in device
memory between steps (see ``State``'s residency protocol). There is no CPU fallback: without the
built library, or without a gfx950 GPU, constructing a Solver raises.
"""

from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

from .config import (
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
from .grid import Bathymetry, Grid, State

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwavecore_hip.so")

WC_OK, WC_EINVAL, WC_EDEVICE, WC_ENOMEM = 0, 1, 2, 3
PRECISION_FP64 = 0
PRECISION_FP32 = 1
PRECOND_JACOBI = 0
PRECOND_NETWORK = 1
BED_FEATURE_SCALE = 10.0

_D = ctypes.c_double
_P = ctypes.POINTER(ctypes.c_double)
_U64 = ctypes.c_uint64
_I32 = ctypes.c_int32
_VP = ctypes.c_void_p

# Every entry point of include/wavecore_hip.h with its ctypes signature.
SIGNATURES = {
    "wc_last_error": (ctypes.c_char_p, []),
    "wc_solver_new": (_I32, [_U64, _U64, _D, _D, _P, ctypes.POINTER(_VP)]),
    "wc_solver_free": (None, [_VP]),
    "wc_solver_set_cfl": (_I32, [_VP, _D]),
    "wc_solver_cfl": (_I32, [_VP, _P]),
    "wc_solver_with_order": (_I32, [_VP, _I32]),
    "wc_solver_with_manning": (_I32, [_VP, _D]),
    "wc_solver_with_sponge": (_I32, [_VP, ctypes.POINTER(CSponge)]),
    "wc_solver_with_wavemaker": (_I32, [_VP, ctypes.POINTER(CWaveMaker)]),
    "wc_solver_with_dispersion": (_I32, [_VP, ctypes.POINTER(CDispersion)]),
    "wc_solver_with_periodic_y": (_I32, [_VP]),
    "wc_solver_with_shallow_cells": (_I32, [_VP, ctypes.POINTER(ctypes.c_uint8), _U64]),
    "wc_solver_bed": (_I32, [_VP, _P]),
    "wc_solver_set_precision": (_I32, [_VP, _I32]),
    "wc_solver_set_option": (_I32, [_VP, ctypes.c_char_p, ctypes.c_int64]),
    "wc_solver_flux_rows": (_I32, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _P, _I32,
                                   ctypes.POINTER(_I32)]),
    "wc_state_set": (_I32, [_VP, _P, _P, _P, _P, _D]),
    "wc_state_get": (_I32, [_VP, _P, _P, _P, _P, _P]),
    "wc_state_set_time": (_I32, [_VP, _D]),
    "wc_stable_dt": (_I32, [_VP, _P]),
    "wc_stable_dt_noalloc": (_I32, [_VP, _VP, _VP, _U64, _U64, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, _VP, _VP]),
    "wc_step": (_I32, [_VP, _D]),
    "wc_limiting_cell": (_I32, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(CLimiting)]),
    "wc_breaking_indicator": (_I32, [_VP, _P]),
    "wc_dispersion_stats": (_I32, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(CDispersionStats)]),
    "wc_state_volume": (_I32, [_VP, _P]),
    "wc_state_max_abs_momentum": (_I32, [_VP, _P]),
    "wc_state_min_depth": (_I32, [_VP, _P]),
    "wc_linear_wave": (_I32, [_D, _D, _I32, ctypes.POINTER(CLinearWave)]),
    "wc_manning_factor": (_D, [_D, _D, _D, _D]),
    "wc_warm_start_weights": (_I32, [_I32, _I32, _P]),
    "wc_check_shared_division": (_I32, [_D, _P, _U64, _P, _P]),
    "wc_stream_ceilings": (_I32, [_U64, _I32, _I32, _P]),
    "wc_run_steps": (_I32, [_VP, _U64, ctypes.POINTER(_U64), _P]),
    "wc_solver_set_preconditioner": (_I32, [_VP, _I32, ctypes.POINTER(ctypes.c_float), _U64]),
    "wc_solver_precond_fallbacks": (_I32, [_VP, ctypes.POINTER(_U64)]),
    "wc_solver_capture": (_I32, [_VP, _I32] + [ctypes.POINTER(ctypes.c_float)] * 5
                          + [ctypes.POINTER(_I32)]),
    "wc_solver_residual_history": (_I32, [_VP, _P, _U64, ctypes.POINTER(_U64)]),
    "wc_stats_new": (_I32, [_VP, _D, _D, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_VP)]),
    "wc_stats_free": (None, [_VP]),
    "wc_stats_record": (_I32, [_VP, _D]),
    "wc_stats_read": (_I32, [_VP, _P, ctypes.POINTER(_U64), _P, _P, _P, _P, _P, _P]),
    "wc_stats_harmonic": (_I32, [_VP, ctypes.POINTER(_I32), _P, _P]),
    "wc_stats_finish": (_I32, [_VP]),
    "wc_watch_init": (_I32, [ctypes.c_void_p, _D]),
    "wc_watch_collapsed": (_I32, [ctypes.c_void_p, _D, _P]),
    "wc_stage_step": (_I32, [_VP, _VP, ctypes.c_void_p, _D, _P, ctypes.POINTER(_I32), _P]),
    "wc_relaxation_new": (_I32, [_VP, ctypes.POINTER(CSponge), _U64, ctypes.POINTER(_VP)]),
    "wc_relaxation_free": (None, [_VP]),
    "wc_relaxation_len": (_U64, [_VP]),
    "wc_relaxation_cells": (_I32, [_VP, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    "wc_relaxation_apply": (_I32, [_VP, _D, _P]),
    "wc_relaxation_apply_device": (_I32, [_VP, _D, _VP]),
    "wc_solver_with_eddy_viscosity": (_I32, [_VP, ctypes.c_void_p]),
    "wc_sampler_new": (_I32, [_VP, _U64, _D, _D, _U64, _U64, _D, _D, _U64,
                              ctypes.POINTER(_VP)]),
    "wc_sampler_free": (None, [_VP]),
    "wc_sampler_sample": (_I32, [_VP, _VP, _I32]),
    "wc_sampler_apply": (_I32, [_VP, _VP, _D, _D]),
    "wc_sampler_read": (_I32, [_VP, _I32, _P, ctypes.POINTER(_D)]),
    "wc_batch_new": (_I32, [ctypes.POINTER(_VP), _U64, ctypes.POINTER(_VP)]),
    "wc_batch_free": (None, [_VP]),
    "wc_batch_set_option": (_I32, [_VP, ctypes.c_char_p, ctypes.c_int64]),
    "wc_batch_launch_widths": (_I32, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "wc_batch_flux_rows": (_I32, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _P, _I32,
                                  ctypes.POINTER(_I32)]),
    "wc_batch_setup_rows": (_I32, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _P, _I32,
                                   ctypes.POINTER(_I32)]),
    "wc_batch_profile_read": (_I32, [_VP, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(_U64)]),
    "wc_batch_run_steps": (_I32, [_VP, _U64, ctypes.POINTER(_U64), _P]),
    "wc_profile_enable": (_I32, [_VP, _I32]),
    "wc_profile_reset": (_I32, [_VP]),
    "wc_profile_read": (_I32, [_VP, ctypes.c_char_p, _U64, _P, ctypes.POINTER(_U64), _U64,
                               ctypes.POINTER(_U64)]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The built library; raises if it is missing (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -m wavesim_amd.build` (hipcc, gfx950)")
        # One HIP runtime per process: torch ships its own libamdhip64.so.7, and the library's
        # dependency of the same name resolves to it if torch is loaded first; the other way
        # round two runtimes would meet and torch would see no device.
        try:
            import torch  # noqa: F401
        except ImportError:  # pragma: no cover - the library then uses ROCm's runtime alone
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class DeviceError(RuntimeError):
    pass


def _check(rc: int) -> None:
    if rc == WC_OK:
        return
    msg = lib().wc_last_error().decode()
    if rc == WC_EINVAL:
        raise ValueError(msg)
    if rc == WC_ENOMEM:
        raise MemoryError(msg)
    raise DeviceError(msg)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


class Solver:
    """wavecore's ``Solver`` on one MI355X GPU (the current HIP device)."""

    def __init__(self, grid: Grid, bed: Bathymetry):
        b = np.ascontiguousarray(bed.b, dtype=np.float64)
        if b.size != grid.cells():
            raise ValueError("bed must be a padded array of grid.cells() values")
        self.grid = grid
        h = _VP()
        _check(lib().wc_solver_new(grid.nx, grid.ny, grid.dx, grid.dy, _ptr(b), ctypes.byref(h)))
        self._h = h
        self.order = Order.Second
        self.manning = 0.0
        self.periodic_y = False
        self._dispersion = None
        self._holding = None  # weak reference to the State on the device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            held = self._holding() if self._holding else None
            if held is not None and not held._host_valid:
                try:
                    self._download(held)
                except Exception:
                    pass
            lib().wc_solver_free(h)
            self._h = None

    # -- builder (solver.rs:301-419) ---------------------------------------------------------
    @property
    def cfl(self) -> float:
        out = ctypes.c_double()
        _check(lib().wc_solver_cfl(self._h, ctypes.byref(out)))
        return out.value

    @cfl.setter
    def cfl(self, v: float) -> None:
        _check(lib().wc_solver_set_cfl(self._h, float(v)))

    def with_order(self, order: Order) -> "Solver":
        _check(lib().wc_solver_with_order(self._h, int(order)))
        self.order = Order(order)
        return self

    def with_manning(self, n: float) -> "Solver":
        _check(lib().wc_solver_with_manning(self._h, float(n)))
        self.manning = float(n)
        return self

    def with_sponge(self, sponge: Sponge) -> "Solver":
        c = c_sponge(sponge)
        _check(lib().wc_solver_with_sponge(self._h, ctypes.byref(c)))
        return self

    def with_wavemaker(self, maker: WaveMaker) -> "Solver":
        c = c_wavemaker(maker)
        _check(lib().wc_solver_with_wavemaker(self._h, ctypes.byref(c)))
        return self

    def with_dispersion(self, d: Dispersion) -> "Solver":
        c = c_dispersion(d)
        _check(lib().wc_solver_with_dispersion(self._h, ctypes.byref(c)))
        if d.breaking.viscosity is not None:
            e = c_eddy(d.breaking.viscosity)
            _check(lib().wc_solver_with_eddy_viscosity(self._h, ctypes.byref(e)))
        self._dispersion = d
        return self

    def with_periodic_y(self) -> "Solver":
        _check(lib().wc_solver_with_periodic_y(self._h))
        self.periodic_y = True
        return self

    def with_shallow_cells(self, cells) -> "Solver":
        a = np.ascontiguousarray(np.asarray(cells, dtype=bool).astype(np.uint8).ravel())
        _check(lib().wc_solver_with_shallow_cells(
            self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), a.size))
        return self

    def set_option(self, key: str, value: int) -> "Solver":
        """Implementation choices that change speed, never results (wc_solver_set_option)."""
        _check(lib().wc_solver_set_option(self._h, key.encode(), int(value)))
        return self

    def flux_rows(self) -> tuple[int, list[tuple[int, float]]]:
        """The marching flux's band height in use (0 before the first flux launch) and the
        tuner's (height, microseconds per launch) measurements (wc_solver_flux_rows)."""
        rows, n = _I32(), _I32()
        cap = 16
        tr = (_I32 * cap)()
        tu = (ctypes.c_double * cap)()
        _check(lib().wc_solver_flux_rows(self._h, ctypes.byref(rows), tr, tu, cap,
                                         ctypes.byref(n)))
        return rows.value, [(tr[q], tu[q]) for q in range(min(n.value, cap))]

    def with_preconditioner(self, params=None) -> "Solver":
        """The dispersive solve's preconditioner: None for the reference's Jacobi diagonal, or
        the learned network's FP32 parameters (wavesim_amd.unet) inside a flexible PCG with the
        same tolerance (the FP32 form; the FP64 path keeps Jacobi)."""
        if params is None:
            _check(lib().wc_solver_set_preconditioner(self._h, PRECOND_JACOBI, None, 0))
            return self
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib().wc_solver_set_preconditioner(
            self._h, PRECOND_NETWORK, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size))
        return self

    def precond_fallbacks(self) -> int:
        """Solves that fell back from the network to Jacobi since the solver was made."""
        out = ctypes.c_uint64()
        _check(lib().wc_solver_precond_fallbacks(self._h, ctypes.byref(out)))
        return int(out.value)

    def capture(self, slot: int = 0) -> dict | None:
        """The last solve's depth, residual r_n and error w* - w_n before iteration n = 0, 2, 5,
        10 (slot 0..3; option "capture" on), as padded float32 grid arrays; None if the solve
        ended before that iteration."""
        n = self.grid.cells()
        out = {k: np.empty(n, dtype=np.float32) for k in ("h", "rx", "ry", "ex", "ey")}
        fp = ctypes.POINTER(ctypes.c_float)
        valid = _I32()
        _check(lib().wc_solver_capture(self._h, slot,
                                       *(out[k].ctypes.data_as(fp)
                                         for k in ("h", "rx", "ry", "ex", "ey")),
                                       ctypes.byref(valid)))
        return out if valid.value else None

    def residual_history(self) -> np.ndarray:
        """r.r before each iteration of the last network-preconditioned solve."""
        out = np.empty(1024)
        n = _U64()
        _check(lib().wc_solver_residual_history(self._h, _ptr(out), 1024, ctypes.byref(n)))
        return out[:n.value].copy()

    def with_precision(self, precision: int) -> "Solver":
        _check(lib().wc_solver_set_precision(self._h, int(precision)))
        return self

    def dispersion(self):
        return self._dispersion

    @property
    def bed(self) -> Bathymetry:
        out = np.empty(self.grid.cells())
        _check(lib().wc_solver_bed(self._h, _ptr(out)))
        return Bathymetry(out)

    # -- residency ---------------------------------------------------------------------------
    def _download(self, s: State) -> None:
        n = self.grid.cells()
        out = [np.empty(n) for _ in range(4)]
        _check(lib().wc_state_get(self._h, *(_ptr(a) for a in out), None))
        s._h, s._hu, s._hv, s._breaking = out

    def _hold(self, s: State) -> None:
        """Make the device hold ``s`` (uploading it if the device copy is not current)."""
        if s._holder is self and s._dev_valid:
            _check(lib().wc_state_set_time(self._h, s.time))
            return
        prev = self._holding() if self._holding else None
        if prev is not None and prev is not s and prev._holder is self:
            prev._sync_host()
            prev._holder = None
            prev._dev_valid = False
        if s._holder is not None and s._holder is not self:
            s._sync_host()
        h, hu, hv, br = (np.ascontiguousarray(a) for a in (s._h, s._hu, s._hv, s._breaking))
        if h.size != self.grid.cells():
            raise ValueError("state does not match the grid")
        _check(lib().wc_state_set(self._h, _ptr(h), _ptr(hu), _ptr(hv), _ptr(br), s.time))
        s._holder = self
        s._dev_valid = True
        s._host_valid = True
        self._holding = weakref.ref(s)

    # -- the API -----------------------------------------------------------------------------
    def stable_dt(self, s: State) -> float:
        """solver.rs:434-450."""
        self._hold(s)
        out = ctypes.c_double()
        _check(lib().wc_stable_dt(self._h, ctypes.byref(out)))
        return out.value

    def step(self, s: State, dt: float) -> None:
        """solver.rs:481-507: advances ``s`` (on the device) and its time by ``dt``."""
        self._hold(s)
        t = s.time
        _check(lib().wc_step(self._h, float(dt)))
        s.time = t + float(dt)
        s._host_valid = False

    def run_steps(self, s: State, steps: int) -> tuple[int, float]:
        """``steps`` iterations of ``dt = stable_dt(s); step(s, dt)`` in one call; stops early
        when dt is not finite and positive. Returns (steps taken, last dt)."""
        self._hold(s)
        done = ctypes.c_uint64()
        last = ctypes.c_double()
        _check(lib().wc_run_steps(self._h, int(steps), ctypes.byref(done), ctypes.byref(last)))
        tt = ctypes.c_double()
        _check(lib().wc_state_get(self._h, None, None, None, None, ctypes.byref(tt)))
        s.time = tt.value
        s._host_valid = False
        return int(done.value), last.value

    def limiting_cell(self, s: State):
        """solver.rs:455-477."""
        self._hold(s)
        found = ctypes.c_int32()
        out = CLimiting()
        _check(lib().wc_limiting_cell(self._h, ctypes.byref(found), ctypes.byref(out)))
        if not found.value:
            return None
        return Limiting(int(out.i), int(out.j), out.depth, out.speed)

    def breaking_indicator(self, s: State) -> np.ndarray:
        """solver.rs:341-349."""
        self._hold(s)
        out = np.empty(self.grid.nx * self.grid.ny)
        _check(lib().wc_breaking_indicator(self._h, _ptr(out)))
        return out

    def dispersion_stats(self):
        """solver.rs:392-396."""
        has = ctypes.c_int32()
        out = CDispersionStats()
        _check(lib().wc_dispersion_stats(self._h, ctypes.byref(has), ctypes.byref(out)))
        if not has.value:
            return None
        return DispersionStats(int(out.solves), int(out.iterations), int(out.unconverged))

    def volume(self, s: State) -> float:
        """State::volume on the device (canonical summation order)."""
        self._hold(s)
        out = ctypes.c_double()
        _check(lib().wc_state_volume(self._h, ctypes.byref(out)))
        return out.value

    def max_abs_momentum(self, s: State) -> float:
        self._hold(s)
        out = ctypes.c_double()
        _check(lib().wc_state_max_abs_momentum(self._h, ctypes.byref(out)))
        return out.value

    def min_depth(self, s: State) -> float:
        self._hold(s)
        out = ctypes.c_double()
        _check(lib().wc_state_min_depth(self._h, ctypes.byref(out)))
        return out.value

    # -- measurement ---------------------------------------------------------------------------
    def profile(self, mode: int) -> None:
        """0: off; 1 (True): events around every launch; 2: events around each batch of PCG
        iterations only (the FP32 form's one kernel: "cg_iter"; the FP64 path's two: "cg_pair")."""
        _check(lib().wc_profile_enable(self._h, int(mode)))

    def profile_reset(self) -> None:
        _check(lib().wc_profile_reset(self._h))

    def profile_read(self) -> dict:
        names = ctypes.create_string_buffer(1024)
        ms = (ctypes.c_double * 64)()
        n = (ctypes.c_uint64 * 64)()
        cnt = ctypes.c_uint64()
        _check(lib().wc_profile_read(self._h, names, 1024, ms, n, 64, ctypes.byref(cnt)))
        keys = names.value.decode().split(";")
        return {k: (ms[q], int(n[q])) for q, k in enumerate(keys[: cnt.value])}


STAGE_OK, STAGE_UNSTABLE, STAGE_NO_WATER, STAGE_COLLAPSED = 0, 1, 2, 3
WATCH_STEPS = 1000
COLLAPSE = 5.0


class CWatch(ctypes.Structure):
    _fields_ = [("reference", ctypes.c_double), ("steps", ctypes.c_uint32),
                ("seconds", ctypes.c_double)]


class StepWatch:
    """wavesim's StepWatch (crates/wavesim/src/lib.rs:849-896): the time step averaged over
    WATCH_STEPS steps, checked for a collapse to a COLLAPSE-th of the still sea's."""

    def __init__(self, reference: float):
        self._c = CWatch()
        _check(lib().wc_watch_init(ctypes.byref(self._c), float(reference)))

    def collapsed(self, stable: float):
        """Counts a step; the ratio at the end of a collapsed block, else None."""
        r = ctypes.c_double()
        _check(lib().wc_watch_collapsed(ctypes.byref(self._c), float(stable), ctypes.byref(r)))
        return r.value if r.value > 0.0 else None


class Stats:
    """wavesim's per-step statistics (Stats and Stage::record, crates/wavesim/src/lib.rs:899-932,
    1276-1300) accumulated on the GPU for the solver's state: recording from time ``from_``,
    waves of ``period`` s; ``elevation`` (nx*ny float32, the caller's bed elevation) or None for
    the solver's bed."""

    def __init__(self, solver: "Solver", from_: float, period: float, elevation=None):
        self.solver = solver
        self.n = solver.grid.nx * solver.grid.ny
        el = None
        if elevation is not None:
            self._el = np.ascontiguousarray(elevation, dtype=np.float32)
            el = self._el.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        h = _VP()
        _check(lib().wc_stats_new(solver._h, float(from_), float(period), el, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().wc_stats_free(self._h)
            self._h = None

    def record(self, dt: float) -> None:
        _check(lib().wc_stats_record(self._h, float(dt)))

    def read(self) -> dict:
        out = {k: np.empty(self.n) for k in ("breaking_seconds", "eta_max", "eta_min",
                                              "wave_mean", "wave_first", "wave_last")}
        sec = ctypes.c_double()
        waves = _U64()
        _check(lib().wc_stats_read(self._h, ctypes.byref(sec), ctypes.byref(waves),
                                   *(_ptr(out[k]) for k in ("breaking_seconds", "eta_max",
                                                            "eta_min", "wave_mean", "wave_first",
                                                            "wave_last"))))
        out["seconds"] = sec.value
        out["waves"] = int(waves.value)
        out["breaking_fraction"] = out["breaking_seconds"] / max(sec.value, 5e-324)
        return out

    def finish(self) -> None:
        """The end of a run (crates/wavesim/src/lib.rs:1451-1453): closes the wave in progress
        if its period has ended."""
        _check(lib().wc_stats_finish(self._h))

    def harmonic(self):
        """FirstHarmonic::amplitudes: (a, -b) per cell, or None."""
        ok = _I32()
        a, b = np.empty(self.n), np.empty(self.n)
        _check(lib().wc_stats_harmonic(self._h, ctypes.byref(ok), _ptr(a), _ptr(b)))
        return (a, b) if ok.value else None


def stage_step(solver: "Solver", s: State, stats: "Stats | None", watch: "StepWatch | None",
               until: float):
    """Stage::next_dt + advance + record (crates/wavesim/src/lib.rs:1385-1434), no frames:
    (dt, status, ratio) with status STAGE_OK when a step of dt was taken."""
    solver._hold(s)
    dt, ratio = ctypes.c_double(), ctypes.c_double()
    status = _I32()
    _check(lib().wc_stage_step(solver._h, stats._h if stats else None,
                               ctypes.byref(watch._c) if watch else None, float(until),
                               ctypes.byref(dt), ctypes.byref(status), ctypes.byref(ratio)))
    if status.value == STAGE_OK:
        tt = ctypes.c_double()
        _check(lib().wc_state_get(solver._h, None, None, None, None, ctypes.byref(tt)))
        s.time = tt.value
        s._host_valid = False
    return dt.value, status.value, ratio.value


def relaxation_targets(target, n: int) -> np.ndarray:
    """(eta, u, v) of each of the n zone cells, interleaved, from the reference's closure
    ``m -> (eta, u, v)`` or from an array of shape (n, 3)."""
    if callable(target):
        out = np.empty((n, 3))
        for m in range(n):
            out[m] = target(m)
    else:
        out = np.array(target, dtype=np.float64).reshape(n, 3)
    return np.ascontiguousarray(out.reshape(-1))


class Sampler:
    """wavesim's coarse-to-fine Sampler (crates/wavesim/src/lib.rs:1548-1616) for a fine
    solver's relaxation zone: ``sample(coarse, state, slot)`` before (slot 0) and after (slot 1)
    the coarse step that brackets the fine step, ``apply(state, end, dt)`` after the fine step
    (targets interpolated in time, then Relaxation::apply)."""

    def __init__(self, zone: "Relaxation", fine_x0: int, fine_grid, coarse_grid, crop: int = 0):
        self.zone = zone
        self.n = len(zone)
        h = _VP()
        _check(lib().wc_sampler_new(zone._r, int(fine_x0), float(fine_grid.dx),
                                    float(fine_grid.dy), coarse_grid.nx, coarse_grid.ny,
                                    float(coarse_grid.dx), float(coarse_grid.dy), int(crop),
                                    ctypes.byref(h)))
        self._s = h

    def __del__(self):
        if getattr(self, "_s", None):
            lib().wc_sampler_free(self._s)
            self._s = None

    def sample(self, coarse: "Solver", state: State, slot: int) -> None:
        coarse._hold(state)
        _check(lib().wc_sampler_sample(self._s, coarse._h, int(slot)))

    def apply(self, state: State, end: float, dt: float) -> None:
        self.zone.solver._hold(state)
        _check(lib().wc_sampler_apply(self._s, self.zone._r, float(end), float(dt)))
        state._host_valid = False

    def read(self, slot: int):
        """(targets as an (n, 3) array, time) of a sample (0, 1) or the last targets (2)."""
        out = np.empty(3 * max(self.n, 1))
        t = ctypes.c_double()
        _check(lib().wc_sampler_read(self._s, int(slot), _ptr(out), ctypes.byref(t)))
        return out[:3 * self.n].reshape(self.n, 3), t.value


class Batch:
    """Many independent solvers stepped together, every phase one launch over all of them
    (wc_batch_*): FP32 form, one-kernel Jacobi PCG. ``run_steps(states, steps)`` advances each
    solver's state ``steps`` times with its own stable dt, as ``Solver.run_steps`` does for one,
    and returns the steps each took and its last dt."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (_VP * len(self.solvers))(*[s._h for s in self.solvers])
        h = _VP()
        _check(lib().wc_batch_new(arr, len(self.solvers), ctypes.byref(h)))
        self._b = h

    def __del__(self):
        if getattr(self, "_b", None):
            lib().wc_batch_free(self._b)
            self._b = None

    def set_option(self, key: str, value: int) -> "Batch":
        _check(lib().wc_batch_set_option(self._b, key.encode(), int(value)))
        return self

    def flux_rows(self) -> tuple[int, list[tuple[int, float]]]:
        """The marching flux's band height (0 with the tile flux) and the tuner's (height,
        microseconds per launch) measurements (wc_batch_flux_rows)."""
        rows, n = _I32(), _I32()
        cap = 16
        tr = (_I32 * cap)()
        tu = (ctypes.c_double * cap)()
        _check(lib().wc_batch_flux_rows(self._b, ctypes.byref(rows), tr, tu, cap,
                                        ctypes.byref(n)))
        return rows.value, [(tr[q], tu[q]) for q in range(min(n.value, cap))]

    def setup_rows(self) -> tuple[int, list[tuple[int, float]]]:
        """The SGN set-up's tile height (64 x rows cells a block) and the tuner's (height,
        microseconds per launch) measurements (wc_batch_setup_rows)."""
        rows, n = _I32(), _I32()
        cap = 8
        tr = (_I32 * cap)()
        tu = (ctypes.c_double * cap)()
        _check(lib().wc_batch_setup_rows(self._b, ctypes.byref(rows), tr, tu, cap,
                                         ctypes.byref(n)))
        return rows.value, [(tr[q], tu[q]) for q in range(min(n.value, cap))]

    def launch_widths(self) -> tuple[int, int]:
        """The persistent widths of the 256- and 1024-thread launches (tuned on the first run)."""
        a, b = _I32(), _I32()
        _check(lib().wc_batch_launch_widths(self._b, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def profile_read(self) -> tuple[float, int]:
        """(milliseconds, launches) of the PCG iteration launches since set_option("profile", 1)."""
        ms, n = ctypes.c_double(), _U64()
        _check(lib().wc_batch_profile_read(self._b, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, int(n.value)

    def run_steps(self, states, steps: int):
        if len(states) != len(self.solvers):
            raise ValueError("one state per solver")
        for g, s in zip(self.solvers, states):
            g._hold(s)
        n = len(self.solvers)
        done = (_U64 * n)()
        last = np.zeros(n)
        _check(lib().wc_batch_run_steps(self._b, int(steps), done, _ptr(last)))
        for g, s in zip(self.solvers, states):
            tt = ctypes.c_double()
            _check(lib().wc_state_get(g._h, None, None, None, None, ctypes.byref(tt)))
            s.time = tt.value
            s._host_valid = False
        return [int(d) for d in done], last


class Relaxation:
    """wavecore's ``Relaxation`` (forcing.rs:227-284): zones where the water is pulled towards
    targets the caller supplies at every step.

    ``Relaxation(solver, zones)`` is ``Relaxation::new(grid, zones)`` for the solver's grid;
    ``apply(state, dt, target)`` is ``apply(grid, bed, state, dt, target)`` with the solver's
    bed, on the device copy of ``state``. ``target`` is the reference's closure
    ``m -> (eta, u, v)``, an (n, 3) array, or a device pointer (an int) to 3n doubles.
    """

    def __init__(self, solver: "Solver", zones):
        self.solver = solver
        arr = (CSponge * max(len(zones), 1))(*[c_sponge(z) for z in zones])
        h = _VP()
        _check(lib().wc_relaxation_new(solver._h, arr, len(zones), ctypes.byref(h)))
        self._r = h

    def __del__(self):
        if getattr(self, "_r", None):
            lib().wc_relaxation_free(self._r)
            self._r = None

    def __len__(self) -> int:
        return int(lib().wc_relaxation_len(self._r))

    def cells(self) -> list[tuple[int, int]]:
        """forcing.rs:252-254: padded (i, j) of each cell, in the order ``apply`` takes."""
        n = len(self)
        i = (_U64 * max(n, 1))()
        j = (_U64 * max(n, 1))()
        _check(lib().wc_relaxation_cells(self._r, i, j))
        return [(int(i[m]), int(j[m])) for m in range(n)]

    def apply(self, s: State, dt: float, target) -> None:
        """forcing.rs:259-283."""
        self.solver._hold(s)
        if isinstance(target, int):
            _check(lib().wc_relaxation_apply_device(self._r, float(dt), _VP(target)))
        else:
            t = relaxation_targets(target, len(self))
            _check(lib().wc_relaxation_apply(self._r, float(dt), _ptr(t)))
        s._host_valid = False


def linear_wave(omega: float, h: float, dispersive: bool) -> LinearWave:
    out = CLinearWave()
    _check(lib().wc_linear_wave(omega, h, int(dispersive), ctypes.byref(out)))
    return LinearWave(out.k, out.phase_speed, out.group_speed)


def manning_factor(h: float, speed: float, dt: float, n: float) -> float:
    return lib().wc_manning_factor(h, speed, dt, n)


def check_shared_division(y: float, x: np.ndarray):
    """x / y on the device both through the flux kernels' shared reciprocal and by the device's
    own division: (shared, plain), which must agree bit for bit."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    shared, plain = np.empty_like(x), np.empty_like(x)
    _check(lib().wc_check_shared_division(float(y), _ptr(x), x.size, _ptr(shared), _ptr(plain)))
    return shared, plain


STREAM_TESTS = ("copy", "copy_nt", "read", "read_nt", "write", "write_nt", "triad", "triad_nt")


def stream_ceilings(bytes_per_array: int = 1 << 30, grid: int = 2048, reps: int = 10) -> dict:
    """What plain streams reach on this GPU's HBM (wc_stream_ceilings): TB/s moved by a copy,
    reads alone, writes alone and three reads with a write, plain and nontemporal."""
    out = np.zeros(8)
    _check(lib().wc_stream_ceilings(int(bytes_per_array), int(grid), int(reps), _ptr(out)))
    return dict(zip(STREAM_TESTS, out.tolist()))


def warm_start_weights(order: int, window: int = 0) -> np.ndarray:
    """The FP32 form's warm-start weights c_1..c_K (host only): exact for solution steps of
    degree < order, of least norm over K = max(window, order) past steps."""
    out = np.zeros(max(int(order), int(window), 1))
    _check(lib().wc_warm_start_weights(int(order), int(window), _ptr(out)))
    return out[:max(int(order), int(window))]
