"""The reference's slow physics checks (waves.rs, tests/dispersion.rs) on the GPU path, in both
arithmetics, with the reference's bounds; each prints what it measured beside the figure the
reference's README publishes. The FP64 path equals the CPU oracle bit for bit, so these runs
also hold the oracle to the same physics at these scenarios.
"""

from __future__ import annotations

import pytest

import verification as ver
from wavesim_amd.solver import PRECISION_FP32, PRECISION_FP64, Solver

pytestmark = pytest.mark.gpu

PRECISIONS = [pytest.param(PRECISION_FP64, id="fp64"), pytest.param(PRECISION_FP32, id="fp32")]


def maker_of(precision):
    return lambda grid, bed: Solver(grid, bed).with_precision(precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_wave_maker_radiates_the_requested_amplitude(precision):
    west, east = ver.check_wave_maker_radiates_the_requested_amplitude(maker_of(precision))
    print(f"maker: west {west:.5f} ({100 * (1 - west / 0.05):.2f}% low), east {east:.5f} "
          f"({100 * (1 - east / 0.05):.2f}% low); reference: 0.4% and 0.6% low")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_shoaling_follows_greens_law(precision):
    measured, expected = ver.check_shoaling_follows_greens_law(maker_of(precision))
    print(f"Green's law: measured {measured:.4f}, expected {expected:.4f} "
          f"({100 * abs(measured / expected - 1):.2f}% off; reference 0.14%)")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_oblique_waves_obey_snells_law_and_wave_action(precision):
    lags, lag0, measured, expected = ver.check_oblique_waves_obey_snells_law(maker_of(precision))
    print(f"Snell: lags {lags[0]:.4f}, {lags[1]:.4f} rad (expected {lag0:.4f}); amplitude ratio "
          f"{measured:.4f} vs {expected:.4f} ({100 * abs(measured / expected - 1):.2f}% off; "
          f"reference 0.02%)")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_friction_attenuates_waves_at_the_predicted_rate(precision):
    att, pred = ver.check_friction_attenuates_waves_at_the_predicted_rate(maker_of(precision))
    print(f"friction: attenuation {att:.4f}, predicted {pred:.4f} (reference 0.919 vs 0.911)")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_solitary_wave_runup_matches_synolakis_law(precision):
    runup, law = ver.check_solitary_wave_runup_matches_synolakis_law(maker_of(precision))
    print(f"runup {runup:.4f} m, Synolakis {law:.4f} m ({100 * (runup / law - 1):+.1f}%; "
          f"reference +1.5%)")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_solitary_wave_keeps_its_shape_and_speed(precision):
    out = ver.check_solitary_wave_keeps_its_shape_and_speed(maker_of(precision))
    for k in ("Second", "Fifth"):
        peak, x, shape = out[k]
        print(f"solitary {k}: peak {peak:.4f}, crest {x:.2f} m (exact {out['expected_x']:.2f}), "
              f"shape error {100 * shape:.2f}%")
    print(f"shallow water: shape error {100 * out['shallow'][2]:.0f}%; reference: amplitude "
          f"0.2001, crest within 0.01 m, shape 0.2%, shallow water 57%")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_wave_maker_with_dispersion(precision):
    amps, lag, expected, its = ver.check_wave_maker_with_dispersion(maker_of(precision))
    print(f"SGN maker: amplitudes {amps[0]:.4f}, {amps[1]:.4f} (reference 0.0498, 0.0499); "
          f"lag {lag:.3f} rad vs SGN {expected:.3f}; {its:.1f} iterations per solve")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_big_swell_keeps_its_energy_in_deep_water(precision):
    near, far, _, _, kept, kept_swe = ver.check_big_swell_keeps_its_energy_in_deep_water(
        maker_of(precision))
    print(f"big swell: equivalent height {near[1]:.2f} -> {far[1]:.2f} m, energy kept "
          f"{100 * kept:.0f}% (shallow water {100 * kept_swe:.0f}%); reference 95% and 14%")


@pytest.mark.parametrize("precision", PRECISIONS)
def test_waves_shoal_then_break_in_a_depth_limited_surf_zone(precision):
    out = ver.check_waves_shoal_then_break_in_a_depth_limited_surf_zone(maker_of(precision))
    print(f"beach: energy height at 6 m {out['energy_height_6m']:.2f} m, shoaling "
          f"x{out['growth']:.2f} (Green x{out['green']:.2f}), tallest H/h "
          f"{out['tallest_h_over_d']:.2f} at {out['tallest_depth']:.2f} m, inner surf H/h up to "
          f"{out['inner_max_h_over_d']:.2f}; reference x1.15, 0.81 at 2.8 m, 1.01")
