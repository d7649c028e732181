"""The FP32 chosen form against the FP64 path, on the GPU.

The FP32 form stores the bulk in FP32 and keeps FP64 where accuracy lives (the state as FP32
pairs with a compensated update, inner products, norms, PCG scalars, surface-minus-bed
differences). It is held to the bounds of its specification: surface elevation within 1e-3 m
of FP64, the breaking mask (indicator >= 0.8, crates/wavesim/src/lib.rs:1222) identical on at
least 99.9% of cells, and the reference's own still-water and conservation checks.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import known_answers as ka
from helpers import bore_state
from wavesim_amd import configs
from wavesim_amd.config import Dispersion, Order
from wavesim_amd.grid import GHOST, Bathymetry, Grid, State
from wavesim_amd.solver import PRECISION_FP32, Solver

pytestmark = pytest.mark.gpu


def fp32(grid, bed, pcg=None):
    """pcg: 3 (two kernels per iteration), 4 (one, 256 blocks marching down column strips),
    512 (one, 512 blocks marching), "tiles" (one, 256 blocks over 64x16 tiles), "walk8" (one,
    blocks over 64x16 tiles taking 8 tiles each or more)."""
    s = Solver(grid, bed).with_precision(PRECISION_FP32)
    if pcg == 512:
        return s.set_option("pcg_kernels", 4).set_option("pcg_lanes", 512)
    if pcg == "tiles":
        return s.set_option("pcg_kernels", 4).set_option("pcg_march", 0)
    if pcg == "walk8":
        return (s.set_option("pcg_kernels", 4).set_option("pcg_march", 0)
                .set_option("pcg_tiles_per_block", 8))
    return s.set_option("pcg_kernels", pcg) if pcg else s


def fp32_with(pcg):
    return lambda grid, bed: fp32(grid, bed, pcg)


def compare(w, s64: State, s32: State, g64: Solver, g32: Solver):
    grid = w.grid
    b = grid.inner(w.bed.b)
    h64, h32 = grid.inner(s64.arrays()[0]), grid.inner(s32.arrays()[0])
    wet = (h64 > 0.05) | (h32 > 0.05)
    eta64, eta32 = h64 + b, h32 + b
    d = np.where(wet, eta32 - eta64, 0.0)
    m64 = g64.breaking_indicator(s64) >= 0.8
    m32 = g32.breaking_indicator(s32) >= 0.8
    return {
        "eta_max_err_m": float(np.max(np.abs(d))),
        "eta_rms_err_m": float(np.sqrt(np.mean(d[wet] ** 2))) if wet.any() else 0.0,
        "mask_agreement": float(np.mean(m64 == m32)),
        "breaking_cells": int(m64.sum()),
        "volume_rel_diff": abs(g32.volume(s32) - g64.volume(s64)) / g64.volume(s64),
    }


def run_pair(w, steps, t0=0.0, state=None, pcg=None):
    g64 = w.solver(Solver)
    g32 = w.solver(fp32_with(pcg))
    s64 = state.clone() if state is not None else w.still_water()
    s64.time = t0
    s32 = s64.clone()
    n64, _ = g64.run_steps(s64, steps)
    n32, _ = g32.run_steps(s32, steps)
    assert n64 == steps and n32 == steps
    return g64, g32, s64, s32


@pytest.mark.parametrize("pcg", [3, 4, 512, "tiles", "walk8"])
def test_fp32_rough_bed_matches_fp64_within_its_bounds(pcg):
    w = configs.c1_rough_bed(nx=256, ny=128)
    g64, g32, s64, s32 = run_pair(w, 600, t0=40.0, pcg=pcg)
    r = compare(w, s64, s32, g64, g32)
    print(f"fp32 (pcg {pcg}) vs fp64:", r, g64.dispersion_stats(), g32.dispersion_stats(),
          flush=True)
    assert r["eta_max_err_m"] < 1e-3, r
    assert r["mask_agreement"] >= 0.999, r
    st = g32.dispersion_stats()
    assert st.unconverged == 0


def test_fp32_rough_bed_over_its_whole_4000_step_run():
    """The rough-bed configuration at its own size and length: 1024x512 cells, the 4000 steps
    of its run from still water, the FP32 form and the FP64 path each stepped alone (no shared
    state after t = 0). Compared every 1000 steps, so that drift shows if it grows: surface
    within 1e-3 m, breaking masks equal on at least 99.9% of cells, volume within 1e-9."""
    w = configs.c1_rough_bed()
    assert (w.grid.nx, w.grid.ny, w.steps) == (1024, 512, 4000)
    g64, g32 = w.solver(Solver), w.solver(fp32_with(None))
    s64 = w.still_water()
    s32 = s64.clone()
    breaking = 0
    for done in range(1000, w.steps + 1, 1000):
        assert g64.run_steps(s64, 1000)[0] == 1000
        assert g32.run_steps(s32, 1000)[0] == 1000
        r = compare(w, s64, s32, g64, g32)
        print(f"step {done}, t {s64.time:.1f} s: {r}", flush=True)
        assert r["eta_max_err_m"] < 1e-3, (done, r)
        assert r["mask_agreement"] >= 0.999, (done, r)
        assert r["volume_rel_diff"] < 1e-9, (done, r)
        assert abs(s32.time - s64.time) < 1e-6 * s64.time
        breaking = max(breaking, r["breaking_cells"])
    st64, st32 = g64.dispersion_stats(), g32.dispersion_stats()
    print(f"iterations per solve: FP64 {st64.iterations / st64.solves:.1f}, FP32 form "
          f"{st32.iterations / st32.solves:.1f}", flush=True)
    assert st64.unconverged == 0 and st32.unconverged == 0
    assert breaking > 0  # the masks are compared where waves break, not on still water alone


@pytest.mark.parametrize("recon", [64, 32])
def test_fp32_large_reef_tile_matches_fp64_within_its_bounds(recon):
    """A tile of the 4096x2048 reef (its slope from the wave-maker towards the reef edge, lobes,
    overlay, Manning, sponges), both flux reconstructions of the FP32 form."""
    w = configs.c4_crop()
    g64 = w.solver(Solver)
    g32 = w.solver(fp32_with(None)).set_option("flux_reconstruction", recon)
    s64 = w.still_water()
    s64.time = 60.0
    s32 = s64.clone()
    assert g64.run_steps(s64, 400)[0] == 400 and g32.run_steps(s32, 400)[0] == 400
    r = compare(w, s64, s32, g64, g32)
    print(f"fp32 (reconstruction {recon}) vs fp64, large reef tile:", r, flush=True)
    assert r["eta_max_err_m"] < 1e-3, r
    assert r["mask_agreement"] >= 0.999, r
    assert g32.dispersion_stats().unconverged == 0


def test_fp32_plane_beach_fv_matches_fp64():
    w = configs.c0_plane_beach()
    g64, g32, s64, s32 = run_pair(w, 2000)
    r = compare(w, s64, s32, g64, g32)
    print("fp32 vs fp64 (FV):", r, flush=True)
    assert r["eta_max_err_m"] < 1e-3, r


@pytest.mark.parametrize("pcg", [3, 4, 512, "tiles", "walk8"])
def test_fp32_lake_at_rest_and_closed_basin_volume(pcg):
    # well_balanced.rs:16-62 and 87-121 on the FP32 form
    for mean_depth, amplitude in [(8.0, 3.0), (0.5, 2.5)]:
        grid = Grid(64, 48, 3.0, 3.0)
        bed = Bathymetry.rough(grid, mean_depth, amplitude)
        s = State.lake_at_rest(grid, bed, 0.0)
        g = fp32(grid, bed, pcg).with_order(Order.Fifth).with_dispersion(Dispersion())
        v0 = s.volume(grid)
        g.run_steps(s, 300)
        assert s.max_abs_momentum(grid) < 1e-9, s.max_abs_momentum(grid)
        assert abs(s.volume(grid) - v0) < 1e-9 * v0
        assert s.min_depth(grid) >= 0.0
    ka.check_walls_and_periodic_seam_keep_volume(fp32_with(pcg),
                                                 cases=[(False, True), (True, True)])


@pytest.mark.parametrize("pcg", [3, 4, 512, "tiles"])
def test_fp32_standing_wave_frequency(pcg):
    out = ka.check_standing_wave_sgn_frequency(fp32_with(pcg), orders=(Order.Fifth,))
    assert abs(out["Fifth"] - 1.4261) < 3e-3, out


def test_switching_precision_converts_the_state():
    w = configs.c1_rough_bed(nx=96, ny=48)
    g = w.solver(Solver)
    s = w.still_water()
    s.time = 30.0
    g.run_steps(s, 50)
    before = [a.copy() for a in s.arrays()[:3]]
    g.with_precision(PRECISION_FP32)
    s._host_valid = False  # fetch the converted state back from the device
    after = s.arrays()[:3]
    for a, b in zip(after, before):
        inner = w.grid.inner(a) - w.grid.inner(b)
        assert np.max(np.abs(inner)) <= 1e-12 * max(1.0, np.max(np.abs(b)))
    n, dt = g.run_steps(s, 20)
    assert n == 20 and math.isfinite(dt)


def test_fp32_breaking_fronts_with_eddy_viscosity_stay_near_fp64():
    """Breaking::viscosity in the FP32 form (the fronts and the viscosity run on the FP64 state
    between the FP32 steps): the surface stays within the form's bound of the FP64 path."""
    from wavesim_amd.config import Breaking, EddyViscosity

    grid = Grid(80, 24, 0.5, 0.5)
    bed = Bathymetry.flat(grid, 1.0)
    disp = Dispersion(breaking=Breaking(viscosity=EddyViscosity()), tolerance=1e-8)
    s = bore_state(grid, 1.0)
    g64 = Solver(grid, bed).with_order(Order.Fifth).with_dispersion(disp)
    g32 = fp32(grid, bed).with_order(Order.Fifth).with_dispersion(disp)
    s64, s32 = s.clone(), s.clone()
    assert g64.run_steps(s64, 160)[0] == 160 and g32.run_steps(s32, 160)[0] == 160
    h64, h32 = grid.inner(s64.arrays()[0]), grid.inner(s32.arrays()[0])
    assert np.max(np.abs(h64 - h32)) < 1e-3
    b64, b32 = s64.arrays()[3], s32.arrays()[3]
    assert np.any(b64 > 0) and np.mean((b64 != 0) == (b32 != 0)) >= 0.99


@pytest.mark.parametrize("recon", [64, 32])
def test_fp32_marching_flux_equals_the_tile_flux_bit_for_bit(recon):
    """The marching flux (one wave per strip) and the 64x4 tile flux reconstruct every cell and
    flux every face with the same arithmetic: the FP32 form steps to the same bits with either,
    over a rough bed with dry land (the first-order fallbacks) and the reef tile's physics."""
    for w in (configs.c1_rough_bed(nx=200, ny=72, steps=60), configs.c4_crop()):
        out = []
        for flux in (2, 3):
            g = w.solver(Solver).with_precision(PRECISION_FP32)
            g.set_option("flux_reconstruction", recon).set_option("flux_kernel", flux)
            s = w.still_water()
            s.time = 30.0
            assert g.run_steps(s, 60)[0] == 60
            out.append([a.copy() for a in s.arrays()[:3]] + [s.time])
        for x, y in zip(out[0][:3], out[1][:3]):
            assert np.array_equal(x, y), w.name
        assert out[0][3] == out[1][3]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_flux_band_height_changes_no_bit_and_the_tuner_keeps_the_fastest(precision):
    """The marching flux's band height (rows per wave: each wave's tile is 58 x rows cells) is a
    speed choice only: fixed heights of 8 rows, 13 (which divides neither grid) and more than
    the grid has step to the same bits as the tuner's choice, in both arithmetics, on a rough
    bed with dry land and on the reef tile. The tuner times its candidates at the first flux
    launch, reports them, and keeps the fastest."""
    for w in (configs.c1_rough_bed(nx=200, ny=72, steps=60), configs.c4_crop()):
        out = []
        for rows in (0, 8, 13, 100000):
            g = w.solver(Solver)
            if precision == "fp32":
                g.with_precision(PRECISION_FP32)
            g.set_option("flux_rows", rows)
            s = w.still_water()
            s.time = 30.0
            assert g.run_steps(s, 40)[0] == 40
            used, tried = g.flux_rows()
            if rows == 0:
                print(f"{w.name} {precision}: tuned band height {used} of {tried}", flush=True)
                assert len(tried) >= 2 and used == min(tried, key=lambda t: t[1])[0]
                assert all(8 <= r <= w.grid.ny and us > 0 for r, us in tried)
            else:
                assert used == min(rows, w.grid.ny) and tried == []
            out.append([a.copy() for a in s.arrays()[:3]] + [s.time])
        for other in out[1:]:
            for x, y in zip(out[0][:3], other[:3]):
                assert np.array_equal(x, y), w.name
            assert out[0][3] == other[3]
