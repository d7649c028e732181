"""The nested run's coarse-to-fine sampler on the GPU (wc_sampler_*): Sampler::sample of the
coarse solver's state before and after a coarse step, the targets interpolated in time, and
the fine zone's relaxation (crates/wavesim/src/lib.rs:1548-1616, 1866-1897), against the
oracle's restatement on the same states, bit for bit (FP64)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from wavesim_amd.config import Order, Sponge, WaveMaker
from wavesim_amd.grid import Bathymetry, Grid, State
from wavesim_amd.solver import Relaxation, Sampler, Solver

pytestmark = pytest.mark.gpu


def coarse_run():
    grid = Grid(64, 24, 6.0, 6.0)
    bed = Bathymetry.plane_beach(grid, 1.0 / 30.0, 8.0)
    g = (Solver(grid, bed).with_order(Order.Second)
         .with_wavemaker(WaveMaker(60.0, 0.5, 10.0, 0.0, 12.0, 1.0)))
    s = State.lake_at_rest(grid, g.bed, 0.0)
    g.run_steps(s, 40)
    return grid, g, s


def test_samples_targets_and_relaxation_match_the_oracle():
    cgrid, coarse, cs = coarse_run()
    crop = 2
    fine_x0 = 30  # fine cells of 2 m: the fine grid starts 60 m into the bed file
    fgrid = Grid(90, 60, 2.0, 2.0)
    fbed = Bathymetry.plane_beach(fgrid, 1.0 / 30.0, 6.0)
    fine = Solver(fgrid, fbed)
    fs = State.lake_at_rest(fgrid, fine.bed, 0.0)
    zone = Relaxation(fine, [Sponge(12, 2.0).edges(True, False, True, True)])
    sm = Sampler(zone, fine_x0, fgrid, cgrid, crop)
    ocs = orc.OracleSampler(zone.cells(), fine_x0, 2.0, 2.0, cgrid.nx, cgrid.ny, 6.0, 6.0, crop)

    def host_sample(s):
        c = s.clone()
        c.fill_walls(cgrid)
        h, hu, hv, _ = c.arrays()
        return ocs.sample(h, hu, hv, coarse.bed.b)

    sm.sample(coarse, cs, 0)
    want0 = host_sample(cs)
    dt = coarse.stable_dt(cs)
    coarse.step(cs, dt)
    sm.sample(coarse, cs, 1)
    want1 = host_sample(cs)
    got0, t0 = sm.read(0)
    got1, t1 = sm.read(1)
    assert t1 == t0 + dt
    assert np.array_equal(got0, want0) and np.array_equal(got1, want1)
    wet = np.isfinite(got1[:, 0])
    assert wet.sum() > 0.5 * len(wet)
    # a fine step ending a third of the way through the coarse one, then the zone's pull
    fdt = dt / 3.0
    end = t0 + fdt
    fine_o = orc.OracleSolver(fgrid, fine.bed)
    zone_o = orc.OracleRelaxation(fine_o, [Sponge(12, 2.0).edges(True, False, True, True)])
    fs.time = t0
    fo = fs.clone()
    sm.apply(fs, end, fdt)
    w = min(max((end - t0) / (t1 - t0), 0.0), 1.0)
    targets = orc.OracleSampler.lerp(want0, want1, w)
    got_t, _ = sm.read(2)
    assert np.array_equal(got_t, targets)
    zone_o.apply(fo, fdt, targets)
    for a, b in zip(fs.arrays()[:3], fo.arrays()[:3]):
        assert np.array_equal(a, b)
    assert np.max(np.abs(fs.arrays()[1])) > 0  # the coarse wave's momentum reached the zone
