"""waveio's run directory (crates/waveio/src/run.rs; docs/data-format.md "Run"), written by
the library's native writer (include/wavesim_run.h) and read back here.

    w = RunWriter.create(dir, bed, ["eta", "breaking"], {"wave_period_s": 14.0})
    w.write_frame(t, [eta, breaking])        # host arrays, like RunWriter::write_frame
    w.frame_from(solver, tide)               # or straight from the GPU state
    w.write_stats(...) / w.stats_from(stats)
    w.finish()                               # run.json, written last
    run = Run.read(dir)

Errors are waveio's: OSError for Error::Io, RunFormatError (a ValueError) for Error::Format and
Error::Json; the library's own argument checks raise ValueError.
"""

from __future__ import annotations

import ctypes
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from .solver import WC_EINVAL, WC_ENOMEM, WC_OK, DeviceError, lib

_D = ctypes.c_double
_U64 = ctypes.c_uint64
_I32 = ctypes.c_int32
_VP = ctypes.c_void_p
_CP = ctypes.c_char_p
_FP = ctypes.POINTER(ctypes.c_float)

WC_EIO, WC_EFORMAT = 4, 5
RUN_FORMAT = "wavesim-run"
RUN_VERSION = 1
RUN_FILES = ("run.json", "bed.f32", "frames.f32", "stats.f32")


class CFrame(ctypes.Structure):
    _fields_ = [("crs", _CP), ("origin_easting", _D), ("origin_northing", _D),
                ("x_bearing_deg", _D), ("note", _CP)]


SIGNATURES = {
    "wc_remove_run": (_I32, [_CP]),
    "wc_run_create": (_I32, [_CP, _CP, _U64, _U64, _D, _D, _FP, ctypes.POINTER(CFrame), _CP,
                             ctypes.POINTER(_CP), _U64, _CP, ctypes.POINTER(_VP)]),
    "wc_run_note": (_I32, [_VP, _CP, _CP]),
    "wc_run_write_frame": (_I32, [_VP, _D, ctypes.POINTER(_FP), _U64, ctypes.POINTER(_U64)]),
    "wc_run_write_stats": (_I32, [_VP, ctypes.POINTER(_CP), ctypes.POINTER(_FP), _U64,
                                  ctypes.POINTER(_U64), _D, _D]),
    "wc_run_finish": (_I32, [_VP, ctypes.c_char_p, _U64]),
    "wc_run_free": (None, [_VP]),
    "wc_run_header": (_I32, [_VP, ctypes.POINTER(_U64), ctypes.c_char_p, _U64,
                             ctypes.POINTER(_U64)]),
    "wc_run_frame_from": (_I32, [_VP, _VP, _D, ctypes.POINTER(_D), ctypes.POINTER(_I32)]),
    "wc_run_stats_from": (_I32, [_VP, _VP]),
}

_bound = False


def rlib() -> ctypes.CDLL:
    global _bound
    L = lib()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


class RunFormatError(ValueError):
    """waveio's Error::Format and Error::Json."""


def _check(rc: int) -> None:
    if rc == WC_OK:
        return
    msg = lib().wc_last_error().decode()
    if rc == WC_EIO:
        raise OSError(msg)
    if rc == WC_EFORMAT:
        raise RunFormatError(msg)
    if rc == WC_EINVAL:
        raise ValueError(msg)
    if rc == WC_ENOMEM:
        raise MemoryError(msg)
    raise DeviceError(msg)


def _json(value) -> bytes | None:
    """JSON text for a value as serde_json::json! would hold it: non-finite floats are null."""
    def clean(v):
        if isinstance(v, float) and not math.isfinite(v):
            return None
        if isinstance(v, dict):
            return {str(k): clean(x) for k, x in v.items()}
        if isinstance(v, (list, tuple)):
            return [clean(x) for x in v]
        if isinstance(v, np.generic):
            return clean(v.item())
        return v
    return None if value is None else json.dumps(clean(value), allow_nan=False).encode()


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32).ravel()


@dataclass
class Frame:
    """Frame (crates/waveio/src/bed.rs:14-24): where the grid sits on the Earth."""
    crs: str = "local"
    origin_easting: float = 0.0
    origin_northing: float = 0.0
    x_bearing_deg: float = 0.0
    note: str = ""


@dataclass
class Bed:
    """The part of waveio's Bed (bed.rs:26-60) a run copies: its name, grid, frame, source
    and f32 elevation (nx * ny, row-major with y outer)."""
    name: str
    nx: int
    ny: int
    dx: float
    dy: float
    elevation: np.ndarray
    frame: Frame = field(default_factory=Frame)
    source: object = None


def remove_run(path) -> None:
    """remove_run (run.rs:56-68)."""
    _check(rlib().wc_remove_run(os.fsencode(path)))


class RunWriter:
    """RunWriter (run.rs:72-213) over the native writer."""

    def __init__(self, handle, bed: Bed, fields):
        self._w = handle
        self.nx, self.ny = bed.nx, bed.ny
        self.fields = list(fields)

    @classmethod
    def create(cls, path, bed: Bed, fields, waves=None) -> "RunWriter":
        """RunWriter::create (run.rs:80-124)."""
        el = _f32(bed.elevation)
        names = (_CP * max(len(fields), 1))(*[f.encode() for f in fields])
        fr = CFrame(bed.frame.crs.encode(), bed.frame.origin_easting, bed.frame.origin_northing,
                    bed.frame.x_bearing_deg, bed.frame.note.encode())
        h = _VP()
        _check(rlib().wc_run_create(os.fsencode(path), bed.name.encode(), bed.nx, bed.ny,
                                    float(bed.dx), float(bed.dy),
                                    el.ctypes.data_as(_FP) if el.size else None,
                                    ctypes.byref(fr), _json(bed.source), names, len(fields),
                                    _json(waves), ctypes.byref(h)))
        return cls(h, bed, fields)

    def __del__(self):
        if getattr(self, "_w", None):
            rlib().wc_run_free(self._w)
            self._w = None

    def _handle(self):
        if not self._w:
            raise ValueError("the run writer has finished")
        return self._w

    def note(self, key: str, value) -> None:
        """RunWriter::note (run.rs:126-132)."""
        _check(rlib().wc_run_note(self._handle(), key.encode(), _json(value) or b"null"))

    def write_frame(self, time: float, fields) -> None:
        """RunWriter::write_frame (run.rs:134-163)."""
        arrs = [_f32(f) for f in fields]
        ptrs = (_FP * max(len(arrs), 1))(*[a.ctypes.data_as(_FP) for a in arrs])
        lens = (_U64 * max(len(arrs), 1))(*[a.size for a in arrs])
        _check(rlib().wc_run_write_frame(self._handle(), float(time), ptrs, len(arrs), lens))

    def write_stats(self, fields, blocks, from_: float, to: float) -> None:
        """RunWriter::write_stats (run.rs:165-196)."""
        arrs = [_f32(b) for b in blocks]
        if len(arrs) != len(fields):
            raise RunFormatError(f"stats must be {len(fields)} blocks of "
                                 f"{self.nx * self.ny} values")
        names = (_CP * max(len(fields), 1))(*[f.encode() for f in fields])
        ptrs = (_FP * max(len(arrs), 1))(*[a.ctypes.data_as(_FP) for a in arrs])
        lens = (_U64 * max(len(arrs), 1))(*[a.size for a in arrs])
        _check(rlib().wc_run_write_stats(self._handle(), names, ptrs, len(arrs), lens,
                                         float(from_), float(to)))

    def frame_from(self, solver, state=None, tide: float = 0.0) -> tuple[bool, float]:
        """Stage::frame_if_due's frame (crates/wavesim/src/lib.rs:1348-1380) from `state` (by
        default the state the solver holds) on the GPU: (written, rise). Not written (the
        reference's Error::Unstable) when the rise is not finite or above 100 m, or an eta is
        not finite."""
        from .solver import STAGE_OK

        if state is not None:
            solver._hold(state)
        rise = _D()
        status = _I32()
        _check(rlib().wc_run_frame_from(self._handle(), solver._h, float(tide),
                                        ctypes.byref(rise), ctypes.byref(status)))
        return status.value == STAGE_OK, rise.value

    def stats_from(self, stats) -> None:
        """Stage::finish's statistics (lib.rs:1436-1509, without the break line)."""
        _check(rlib().wc_run_stats_from(self._handle(), stats._h))

    def header(self) -> tuple[int, str]:
        """(frame_count, the text run.json would hold now)."""
        n, need = _U64(), _U64()
        _check(rlib().wc_run_header(self._handle(), ctypes.byref(n), None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(rlib().wc_run_header(self._handle(), None, buf, need.value, None))
        return n.value, buf.value.decode()

    def finish(self) -> str:
        """RunWriter::finish (run.rs:198-212): the path of run.json."""
        h, self._w = self._handle(), None
        buf = ctypes.create_string_buffer(4096)
        _check(rlib().wc_run_finish(h, buf, len(buf)))
        return os.fsdecode(buf.value)


def read_header(path) -> dict:
    """RunHeader::read (run.rs:224-238)."""
    p = os.path.join(path, "run.json")
    with open(p, encoding="utf-8") as f:
        try:
            h = json.load(f)
        except json.JSONDecodeError as e:
            raise RunFormatError(f"{p}: invalid JSON: {e}") from None
    if h.get("format") != RUN_FORMAT or h.get("version") != RUN_VERSION:
        raise RunFormatError(f"{p}: not a wavesim-run version 1")
    return h


class Run:
    """Run (run.rs:215-296): a finished run read back from disk."""

    def __init__(self, header: dict, data: np.ndarray, stats: np.ndarray | None):
        self.header, self.data, self.stats = header, data, stats

    @classmethod
    def read(cls, path) -> "Run":
        h = read_header(path)

        def floats(p, want):
            with open(p, "rb") as f:
                b = f.read()
            if len(b) != want * 4:
                raise RunFormatError(f"{p}: {len(b)} bytes, header promises {want * 4}")
            return np.frombuffer(b, dtype="<f4")

        n = h["nx"] * h["ny"]
        data = floats(os.path.join(path, h["frames"]), h["frame_count"] * len(h["fields"]) * n)
        st = h.get("stats")
        stats = floats(os.path.join(path, st["file"]), len(st["fields"]) * n) if st else None
        return cls(h, data, stats)

    def stat(self, name: str):
        st = self.header.get("stats")
        if not st or name not in st["fields"] or self.stats is None:
            return None
        n = self.header["nx"] * self.header["ny"]
        f = st["fields"].index(name)
        return self.stats[f * n:(f + 1) * n]

    def field(self, name: str, k: int):
        if name not in self.header["fields"]:
            return None
        n = self.header["nx"] * self.header["ny"]
        start = (k * len(self.header["fields"]) + self.header["fields"].index(name)) * n
        return self.data[start:start + n]

    def frame(self, k: int):
        return self.field("eta", k)
