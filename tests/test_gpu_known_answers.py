"""The reference's known-answer tests on the GPU path (through the C ABI), with the
reference's own bounds. The same checks run on the CPU oracle in test_oracle_known_answers.py."""

from __future__ import annotations

import pytest

import known_answers as ka
from wavesim_amd.solver import PRECISION_FP32, PRECISION_FP64, Relaxation, Solver, linear_wave

pytestmark = pytest.mark.gpu


def make(grid, bed):
    return Solver(grid, bed)


def test_lake_at_rest_over_rough_wet_bed():
    ka.check_lake_at_rest_over_rough_wet_bed(make)


def test_lake_at_rest_with_dry_bumps():
    ka.check_lake_at_rest_with_dry_bumps(make)


def test_dam_break_conserves_volume_and_stays_positive():
    ka.check_dam_break_conserves_volume_and_stays_positive(make)


def test_walls_and_a_periodic_seam_keep_the_water():
    ka.check_walls_and_periodic_seam_keep_volume(make)


def test_a_cell_that_is_no_longer_a_number_is_seen_not_skipped():
    ka.check_nan_cell_is_seen_not_skipped(make)


def test_the_limiting_cell_is_the_fastest_water():
    ka.check_limiting_cell_is_fastest_water(make)


def test_still_water_is_limited_by_its_depth_and_dry_land_by_nothing():
    ka.check_still_water_limited_by_depth_dry_land_by_nothing(make)


def test_the_linear_wave_relation_inverts_the_dispersion_relation():
    ka.check_linear_wave_inverts_dispersion_relation(linear_wave)


def test_a_standing_wave_oscillates_at_the_sgn_frequency():
    out = ka.check_standing_wave_sgn_frequency(make)
    assert abs(out["Fifth"] - 1.4261) < 2e-3, out


def test_still_water_stays_still_with_dispersion():
    ka.check_still_water_with_dispersion(make)


def test_a_swell_crosses_a_kilometre_of_shelf_on_coarse_cells():
    kept = ka.check_swell_crosses_shelf_on_coarse_cells(make)
    assert abs(kept - 1.007) < 0.01, kept


def test_wet_bed_dam_break_matches_stoker():
    ka.check_dam_break(make, 2.0, 0.5)


def test_dry_bed_dam_break_matches_ritter():
    ka.check_dam_break(make, 2.0, 0.0)


def test_a_sponge_over_land_never_speeds_the_water_up():
    ka.check_sponge_over_land_never_speeds_water_up(make)


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_FP32])
def test_relaxation_known_answers(precision):
    def mk(grid, bed):
        return Solver(grid, bed).with_precision(precision)

    ka.check_relaxing_towards_still_water_keeps_a_lake_at_rest(mk, Relaxation)
    ka.check_relaxation_moves_depth_and_momentum_together(mk, Relaxation)
    ka.check_relaxation_zone_cells_and_rates(mk, Relaxation)


def test_held_cells_run_exactly_as_shallow_water():
    ka.check_held_cells_run_as_shallow_water(make)


def test_only_the_held_cells_change():
    ka.check_only_held_cells_change(make)


def test_the_breaking_indicator_reads_the_surface_not_the_held_cells():
    ka.check_breaking_indicator_reads_surface_not_held_cells(make)


def test_a_mask_of_the_wrong_size_is_refused():
    ka.check_wrong_size_mask_refused(make)
