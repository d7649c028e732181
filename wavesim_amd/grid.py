"""Grid, bed and state: the arrays that cross wavecore's API, as numpy arrays on the host.

Mirrors crates/wavecore/src/grid.rs, bathymetry.rs and state.rs: padded row-major arrays with
``GHOST`` ghost layers, ``k = j * width + i``, interior ``i, j`` in ``[GHOST, n + GHOST)``.
The ghost fills keep the reference's loop order, so grids narrower than three cells fill
exactly as the reference does.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Iterator

import numpy as np

GHOST = 3
G = 9.81
H_DRY = 1e-8
THREADED_CELLS = 40_000  # crates/wavecore/src/lib.rs:56 (par.rs:13)


@dataclass(frozen=True)
class Grid:
    """Uniform Cartesian grid in metres (grid.rs:12-58)."""

    nx: int
    ny: int
    dx: float
    dy: float

    def __post_init__(self) -> None:
        if not (self.nx > 0 and self.ny > 0):
            raise ValueError("grid must have at least one interior cell")
        if not (self.dx > 0.0 and self.dy > 0.0):
            raise ValueError("cell size must be positive")
        object.__setattr__(self, "nx", int(self.nx))
        object.__setattr__(self, "ny", int(self.ny))
        object.__setattr__(self, "dx", float(self.dx))
        object.__setattr__(self, "dy", float(self.dy))

    def width(self) -> int:
        return self.nx + 2 * GHOST

    def rows(self) -> int:
        return self.ny + 2 * GHOST

    def cells(self) -> int:
        return self.width() * self.rows()

    def idx(self, i: int, j: int) -> int:
        return j * self.width() + i

    def interior(self) -> Iterator[tuple[int, int]]:
        for j in range(GHOST, GHOST + self.ny):
            for i in range(GHOST, GHOST + self.nx):
                yield i, j

    def centre(self, i: int, j: int) -> tuple[float, float]:
        return ((i - GHOST) + 0.5) * self.dx, ((j - GHOST) + 0.5) * self.dy

    def centres(self) -> tuple[np.ndarray, np.ndarray]:
        """Cell-centre x of every padded column and y of every padded row (ghosts included)."""
        x = (np.arange(self.width(), dtype=np.float64) - GHOST + 0.5) * self.dx
        y = (np.arange(self.rows(), dtype=np.float64) - GHOST + 0.5) * self.dy
        return x, y

    def view(self, a: np.ndarray) -> np.ndarray:
        """A padded flat array seen as [row j, column i]."""
        return a.reshape(self.rows(), self.width())

    def inner(self, a: np.ndarray) -> np.ndarray:
        """The interior of a padded flat array, shape (ny, nx)."""
        return self.view(a)[GHOST : GHOST + self.ny, GHOST : GHOST + self.nx]


def mirror_ghosts(grid: Grid, a: np.ndarray) -> None:
    """Mirror the interior into every ghost layer (grid.rs:61-75), in place."""
    v = grid.view(a)
    nx, ny = grid.nx, grid.ny
    rows = slice(GHOST, GHOST + ny)
    for k in range(GHOST):
        v[rows, GHOST - 1 - k] = v[rows, GHOST + k]
        v[rows, GHOST + nx + k] = v[rows, GHOST + nx - 1 - k]
    for k in range(GHOST):
        v[GHOST - 1 - k, :] = v[GHOST + k, :]
        v[GHOST + ny + k, :] = v[GHOST + ny - 1 - k, :]


def wrap_ghosts_y(grid: Grid, a: np.ndarray) -> None:
    """Wrap the south and north ghost rows around the interior (grid.rs:79-87), in place."""
    v = grid.view(a)
    ny = grid.ny
    for k in range(GHOST):
        v[GHOST - 1 - k, :] = v[GHOST + ny - 1 - k, :]
        v[GHOST + ny + k, :] = v[GHOST + k, :]


def flip_normals(grid: Grid, vx: np.ndarray, vy: np.ndarray) -> None:
    """Negate x components in the x ghost columns, y components in the y ghost rows."""
    a, b = grid.view(vx), grid.view(vy)
    nx, ny = grid.nx, grid.ny
    for k in range(GHOST):
        a[:, GHOST - 1 - k] = -a[:, GHOST - 1 - k]
        a[:, GHOST + nx + k] = -a[:, GHOST + nx + k]
    for k in range(GHOST):
        b[GHOST - 1 - k, :] = -b[GHOST - 1 - k, :]
        b[GHOST + ny + k, :] = -b[GHOST + ny + k, :]


class Bathymetry:
    """Padded bed elevation ``b`` (negative underwater), ghosts mirrored (bathymetry.rs:6-62)."""

    def __init__(self, b: np.ndarray):
        self.b = np.ascontiguousarray(b, dtype=np.float64)

    def copy(self) -> "Bathymetry":
        return Bathymetry(self.b.copy())

    @staticmethod
    def from_fn(grid: Grid, f: Callable[[np.ndarray, np.ndarray], np.ndarray]) -> "Bathymetry":
        """``f(x, y)`` is called once on 2-D arrays of interior cell centres (ny, nx)."""
        x, y = grid.centres()
        xi = x[GHOST : GHOST + grid.nx]
        yj = y[GHOST : GHOST + grid.ny]
        X, Y = np.meshgrid(xi, yj)
        b = np.zeros(grid.cells())
        grid.inner(b)[:, :] = np.broadcast_to(np.asarray(f(X, Y), dtype=np.float64), X.shape)
        mirror_ghosts(grid, b)
        return Bathymetry(b)

    @staticmethod
    def from_interior(grid: Grid, values: np.ndarray) -> "Bathymetry":
        values = np.asarray(values, dtype=np.float64).ravel()
        if values.size != grid.nx * grid.ny:
            raise ValueError(f"bed has {values.size} values, grid needs {grid.nx * grid.ny}")
        b = np.zeros(grid.cells())
        grid.inner(b)[:, :] = values.reshape(grid.ny, grid.nx)
        mirror_ghosts(grid, b)
        return Bathymetry(b)

    @staticmethod
    def flat(grid: Grid, depth: float) -> "Bathymetry":
        return Bathymetry.from_fn(grid, lambda x, y: np.full_like(x, -depth))

    @staticmethod
    def plane_beach(grid: Grid, slope: float, offshore_depth: float) -> "Bathymetry":
        return Bathymetry.from_fn(grid, lambda x, y: -offshore_depth + slope * x)

    @staticmethod
    def rough(grid: Grid, mean_depth: float, amplitude: float) -> "Bathymetry":
        """Deterministic two-mode rough bed (bathymetry.rs:53-62)."""
        lx = grid.nx * grid.dx
        ly = grid.ny * grid.dy
        tau = 2.0 * math.pi

        def f(x, y):
            s = 0.6 * np.sin(tau * 3.0 * x / lx) * np.cos(tau * 2.0 * y / ly) + 0.4 * np.sin(
                tau * 7.0 * x / lx + 1.3
            )
            return -mean_depth + amplitude * s

        return Bathymetry.from_fn(grid, f)


class State:
    """Conserved variables ``h, hu, hv``, the breaking memory and the time (state.rs:7-17).

    The arrays are host numpy arrays. A GPU solver may hold the newest values on the device;
    reading an array property fetches them first, and then counts the host copy as possibly
    modified, so the next step sends it back. A stepping loop that only calls ``stable_dt`` and
    ``step`` never copies anything.
    """

    def __init__(self, h, hu, hv, breaking=None, time: float = 0.0):
        self._h = np.ascontiguousarray(h, dtype=np.float64)
        self._hu = np.ascontiguousarray(hu, dtype=np.float64)
        self._hv = np.ascontiguousarray(hv, dtype=np.float64)
        self._breaking = (
            np.zeros_like(self._h)
            if breaking is None
            else np.ascontiguousarray(breaking, dtype=np.float64)
        )
        self.time = float(time)
        self._holder = None  # the solver whose device holds this state
        self._host_valid = True  # host arrays hold the newest values
        self._dev_valid = False  # the holder's device copy holds the newest values

    # -- residency protocol used by device-backed solvers --------------------------------
    def _sync_host(self) -> None:
        if not self._host_valid:
            self._holder._download(self)
            self._host_valid = True

    def _touch(self) -> np.ndarray:
        self._sync_host()
        self._dev_valid = False

    @property
    def h(self) -> np.ndarray:
        self._touch()
        return self._h

    @property
    def hu(self) -> np.ndarray:
        self._touch()
        return self._hu

    @property
    def hv(self) -> np.ndarray:
        self._touch()
        return self._hv

    @property
    def breaking(self) -> np.ndarray:
        self._touch()
        return self._breaking

    @h.setter
    def h(self, v):
        self._touch()
        self._h = np.ascontiguousarray(v, dtype=np.float64)

    @hu.setter
    def hu(self, v):
        self._touch()
        self._hu = np.ascontiguousarray(v, dtype=np.float64)

    @hv.setter
    def hv(self, v):
        self._touch()
        self._hv = np.ascontiguousarray(v, dtype=np.float64)

    def arrays(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        """``(h, hu, hv, breaking)`` read-only snapshot, without invalidating the device copy."""
        self._sync_host()
        out = []
        for a in (self._h, self._hu, self._hv, self._breaking):
            c = a.view()
            c.flags.writeable = False
            out.append(c)
        return tuple(out)

    def clone(self) -> "State":
        h, hu, hv, br = self.arrays()
        return State(h.copy(), hu.copy(), hv.copy(), br.copy(), self.time)

    # -- constructors and diagnostics (state.rs) ------------------------------------------
    @staticmethod
    def zeros(grid: Grid) -> "State":
        n = grid.cells()
        return State(np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n), 0.0)

    @staticmethod
    def lake_at_rest(grid: Grid, bed: Bathymetry, eta: float) -> "State":
        s = State.zeros(grid)
        grid.inner(s._h)[:, :] = np.maximum(eta - grid.inner(bed.b), 0.0)
        s.fill_walls(grid)
        return s

    def fill_walls(self, grid: Grid) -> None:
        self._touch()
        for a in (self._h, self._hu, self._hv, self._breaking):
            mirror_ghosts(grid, a)
        flip_normals(grid, self._hu, self._hv)

    def fill_boundaries(self, grid: Grid, periodic_y: bool) -> None:
        self.fill_walls(grid)
        if periodic_y:
            for a in (self._h, self._hu, self._hv, self._breaking):
                wrap_ghosts_y(grid, a)

    def volume(self, grid: Grid) -> float:
        """Serial interior sum times dx dy, in the reference's order (state.rs:79-85)."""
        h = self.arrays()[0]
        total = -0.0
        for v in grid.inner(h).ravel().tolist():
            total += v
        return total * grid.dx * grid.dy

    def max_abs_momentum(self, grid: Grid) -> float:
        _, hu, hv, _ = self.arrays()
        # fold(0.0, f64::max): NaN cells drop out, as in the reference
        m = np.fmax(np.abs(grid.inner(hu)), np.abs(grid.inner(hv)))
        return float(np.fmax.reduce(m.ravel(), initial=0.0))

    def min_depth(self, grid: Grid) -> float:
        h = self.arrays()[0]
        return float(np.fmin.reduce(grid.inner(h).ravel(), initial=math.inf))
