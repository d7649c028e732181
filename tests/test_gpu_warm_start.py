"""The FP32 form's extrapolated warm start on the GPU, against the oracle's restatement.

The start of each SGN solve is extrapolated from the last solutions (WarmStart in
wavesim_amd/csrc/wc_kernels.h; oracle: ws_extrapolate). Solves still stop on the reference's
residual test, so the run keeps its waves; what changes is the iteration count, which must
follow the oracle's with the same warm start (a wrong ring slot or coefficient would start the
solves far off and cost iterations).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from wavesim_amd import configs
from wavesim_amd.solver import PRECISION_FP32, Solver

pytestmark = pytest.mark.gpu


def run(solver, w, steps, t0=40.0):
    s = w.still_water()
    s.time = t0  # mid-ramp: the maker drives the dispersion
    for _ in range(steps):
        solver.step(s, solver.stable_dt(s))
    st = solver.dispersion_stats()
    return s, st.iterations / st.solves, st.unconverged


def test_iterations_follow_the_oracle_and_the_waves_stay():
    """Each warm start's iteration count on the GPU lies between the oracle's (FP64) with the
    same weights and the oracle's at the order below: the FP32 form's rounding costs it a little
    against FP64 (most for weights whose squares add up most, which carry the most of the past
    solutions' rounding into the start), a wrong ring slot or weight would cost it much more."""
    w = configs.c1_rough_bed(nx=160, ny=64, steps=80)
    oracle_it = {}
    for order, window in ((0, 0), (1, 1), (2, 2), (3, 3), (3, 6)):
        o = w.solver(lambda gr, b: orc.OracleSolver(gr, b)).with_warm_start(order, window)
        so, oracle_it[(order, window)], uo = run(o, w, w.steps)
        assert uo == 0
        if order == 0:
            continue
        g = (w.solver(Solver).with_precision(PRECISION_FP32).set_option("warm_start", order)
             .set_option("warm_start_window", window))
        sg, itg, ug = run(g, w, w.steps)
        assert ug == 0
        below = oracle_it[(order - 1, order - 1)]
        print(f"order {order} window {window}: GPU FP32 {itg:.2f} iterations per solve, oracle "
              f"FP64 {oracle_it[(order, window)]:.2f} (order {order - 1}: {below:.2f})",
              flush=True)
        assert oracle_it[(order, window)] - 1.0 <= itg < below
        dh = np.abs(w.grid.inner(sg.arrays()[0]) - w.grid.inner(so.arrays()[0]))
        assert float(np.max(dh)) < 1e-4


def test_fewer_iterations_than_the_reference_start_and_the_same_waves():
    w = configs.c1_rough_bed(nx=160, ny=64, steps=80)
    runs = {}
    for order in (0, 3):  # the default window
        g = w.solver(Solver).with_precision(PRECISION_FP32).set_option("warm_start", order)
        runs[order] = run(g, w, w.steps)
    (s0, it0, u0), (s3, it3, u3) = runs[0], runs[3]
    print(f"reference start {it0:.2f}, extrapolated {it3:.2f} iterations per solve", flush=True)
    assert u0 == 0 and u3 == 0
    assert it3 < 0.7 * it0
    dh = np.abs(w.grid.inner(s3.arrays()[0]) - w.grid.inner(s0.arrays()[0]))
    assert float(np.max(dh)) < 1e-4


def test_the_order_is_checked():
    w = configs.c1_rough_bed(nx=32, ny=16, steps=1)
    g = w.solver(Solver).with_precision(PRECISION_FP32)
    with pytest.raises(ValueError):
        g.set_option("warm_start", 6)
    with pytest.raises(ValueError):
        g.set_option("warm_start_window", 13)
    with pytest.raises(ValueError):
        g.set_option("warm_start", -1)
