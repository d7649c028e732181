"""The batched spot-sheet sweep's reef sheets (wavesim_amd.configs.c3_reef_instance) against the
oracle and the FP64 path.

Each size class's first draw (its swell angle, tide, lobes, channel, H and T) runs on a 256x128
sheet on the FP64 path in lockstep with the oracle, bit for bit; and sheets of three sizes up to
the sweep's largest, 1536x768, stepped together in one batch in the FP32 form, stay within the
form's bounds of the FP64 path stepped alone: surface within 1e-3 m, breaking masks equal on at
least 99.9% of cells."""

from __future__ import annotations

import numpy as np
import pytest

from test_gpu_parity import lockstep
from wavesim_amd import configs
from wavesim_amd.grid import State
from wavesim_amd.solver import PRECISION_FP32, Batch, Solver

pytestmark = pytest.mark.gpu

# the first seed of the sweep (3000, 3001, ...) that draws each size
FIRST_OF_SIZE = {(256, 128): 3000, (512, 256): 3012, (768, 384): 3001, (1024, 512): 3002,
                 (1536, 768): 3005}


def started(w, periods=2.0):
    """Still water at t = `periods` wave periods: the maker part way up its ramp (2), or at full
    strength from the first step (4, the end of its ramp)."""
    s = w.still_water()
    s.time = periods * w.notes["T_s"]
    return s


@pytest.mark.parametrize("size", sorted(FIRST_OF_SIZE))
def test_reef_sheet_draw_is_bitwise_the_oracle(size):
    seed = FIRST_OF_SIZE[size]
    w = configs.c3_reef_instance(seed, size=(256, 128))
    assert configs.c3_reef_instance(seed).grid.nx == size[0]
    lockstep(w.grid, w.bed, w.configure, started(w), 120, f"reef sheet {seed} ({size} draw)")


def test_batched_fp32_sheets_stay_within_the_forms_bounds_of_fp64():
    # the sweep's draws with the largest swell at three sizes (H 2.84, 2.91 and 2.95 m)
    ws = [configs.c3_reef_instance(seed) for seed in (3023, 3021, 3020)]
    assert [(w.grid.nx, w.grid.ny) for w in ws] == [(256, 128), (512, 256), (1536, 768)]
    steps = 1500  # from the end of the ramp, long enough for the waves to break on the reefs
    g32 = [w.solver(Solver).with_precision(PRECISION_FP32).set_option("pcg_tiles_per_block", 8)
           for w in ws]
    s32 = [started(w, 4.0) for w in ws]
    done, _ = Batch(g32).run_steps(s32, steps)
    assert done == [steps] * len(ws)
    breaking, top = 0, 0.0
    for w, g, s in zip(ws, g32, s32):
        g64 = w.solver(Solver)
        s64 = started(w, 4.0)
        assert g64.run_steps(s64, steps)[0] == steps
        grid = w.grid
        b = grid.inner(w.bed.b)
        h64, h32 = grid.inner(s64.arrays()[0]), grid.inner(s.arrays()[0])
        wet = (h64 > 0.05) | (h32 > 0.05)
        err = float(np.max(np.abs(np.where(wet, (h32 + b) - (h64 + b), 0.0))))
        i64, i32 = g64.breaking_indicator(s64), g.breaking_indicator(s)
        m64, m32 = i64 >= 0.8, i32 >= 0.8
        agree = float(np.mean(m64 == m32))
        top = max(top, float(np.max(i64)))
        print(f"{w.name}: breaking indicator up to {float(np.max(i64)):.3f}, FP32 against FP64 "
              f"within {float(np.max(np.abs(i32 - i64))):.2e}", flush=True)
        breaking += int(m64.sum())
        print(f"{w.name}: t {s64.time:.1f} s, surface error {err:.2e} m, masks agree {agree:.6f} "
              f"({int(m64.sum())} breaking), iterations per solve FP32 "
              f"{g.dispersion_stats().iterations / g.dispersion_stats().solves:.1f}, FP64 "
              f"{g64.dispersion_stats().iterations / g64.dispersion_stats().solves:.1f}",
              flush=True)
        assert err <= 1e-3, (w.name, err)
        assert agree >= 0.999, (w.name, agree)
        assert abs(s.time - s64.time) < 1e-6 * s64.time
    assert breaking > 0  # the masks are not compared on still water alone
