"""The run's per-step statistics on the GPU (wc_stats_*, wc_stage_step) against the oracle's
restatement of Stage::record (crates/wavesim/src/lib.rs:1276-1300), in lockstep on the FP64 path:
every per-cell field, the wave count and the first-harmonic amplitudes must be equal bit for bit
(the oracle runs in the GPU's canonical summation order with the portable libm)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as orc
from wavesim_amd import configs
from wavesim_amd.solver import STAGE_OK, Solver, Stats, StepWatch, stage_step

pytestmark = pytest.mark.gpu


def test_stage_statistics_match_the_oracle():
    w = configs.c1_rough_bed(nx=96, ny=48, seed=11)
    g = w.solver(Solver)
    o = w.solver(lambda gr, b: orc.OracleSolver(gr, b, dot_order=orc.DOT_CANONICAL16,
                                                math=orc.MATH_PORTABLE))
    s = w.still_water()
    s.time = 20.0
    so = s.clone()
    period = 14.0
    elevation = w.grid.inner(w.bed.b).astype(np.float32).ravel() + np.float32(0.25)
    sg = Stats(g, 22.0, period, elevation)
    st = orc.OracleStats(o, 22.0, period, elevation)
    watch = StepWatch(g.stable_dt(s))
    until = 20.0 + 3.2 * period
    n = 0
    while s.time < until - 1e-9:
        dt, status, _ = stage_step(g, s, sg, watch, until)
        assert status == STAGE_OK
        do = min(o.stable_dt(so), until - so.time)
        assert dt == do
        o.step(so, do)
        st.record(so, do)
        n += 1
    assert n > 100
    rg, ro = sg.read(), st.read()
    assert rg["waves"] == ro["waves"] == 3
    assert rg["seconds"] == ro["seconds"]
    for k in ("breaking_seconds", "eta_max", "eta_min", "wave_mean", "wave_first",
              "wave_last"):
        assert np.array_equal(rg[k], ro[k]), k
    hg, ho = sg.harmonic(), st.harmonic()
    assert hg is not None and ho is not None
    assert np.array_equal(hg[0], ho[0]) and np.array_equal(hg[1], ho[1])
    assert np.max(rg["wave_mean"]) > 0.1  # the waves did arrive
    print("breaking cells", int(np.sum(rg["breaking_seconds"] > 0)), "max wave",
          float(np.max(rg["wave_mean"])), flush=True)


def test_stage_step_reports_what_the_reference_treats_as_errors():
    from wavesim_amd.grid import Bathymetry, Grid, State
    from wavesim_amd.solver import STAGE_NO_WATER

    grid = Grid(8, 8, 1.0, 1.0)
    bed = Bathymetry.flat(grid, -1.0)  # dry land everywhere
    g = Solver(grid, bed)
    s = State.lake_at_rest(grid, bed, 0.0)
    dt, status, _ = stage_step(g, s, None, None, 10.0)
    assert status == STAGE_NO_WATER and math.isinf(dt)


def test_statistics_outliving_their_solver_fail_cleanly():
    """Freeing the solver first (as a garbage collector may) leaves the statistics and the
    relaxation zones detached: their calls raise, their own free is safe."""
    import gc

    from wavesim_amd.config import Sponge
    from wavesim_amd.grid import Bathymetry, Grid
    from wavesim_amd.solver import Relaxation, _check, _ptr, lib

    grid = Grid(16, 8, 1.0, 1.0)
    bed = Bathymetry.flat(grid, 2.0)
    g = Solver(grid, bed)
    st = Stats(g, 0.0, 5.0)
    rl = Relaxation(g, [Sponge(3, 0.5).edges(True, False, False, False)])
    assert len(rl) == 3 * 8
    Solver.__del__(g)  # the solver's own free, while the Python objects still exist
    with pytest.raises(ValueError, match="freed"):
        st.record(0.1)
    with pytest.raises(ValueError, match="freed"):
        st.harmonic()
    targets = np.zeros(3 * len(rl))
    with pytest.raises(ValueError, match="freed"):
        _check(lib().wc_relaxation_apply(rl._r, 0.1, _ptr(targets)))
    del st, rl
    gc.collect()
