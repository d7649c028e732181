"""The FP32 form's extrapolated warm start of the SGN solves, on the oracle.

The reference starts each PCG solve from the previous solution (dispersion.rs:220-222, 488-490).
The FP32 form starts solve n from x_{n-1} + sum_{j=1..m} (-1)^(j+1) C(m, j) d_{n-jP}, with
d_n = x_n - x_{n-1} and P solves per step (2 for SSP-RK2, 1 for Euler). Pinned here by what
the mathematics fixes: exact when the solutions are a polynomial of degree <= m in time plus,
per kind of solve (RK stage), one of degree <= m - 1; not exact one degree higher in either
part; error falling as dt^(m+1) on smooth signals; order 0 = the reference's start; and on a
whole run, the same waves (both solves meet the same residual test) in fewer iterations.
"""

from __future__ import annotations

import math
import os

import numpy as np
import pytest

from oracle import oracle as orc
from wavesim_amd import configs


def history(fn, n, period):
    """past[i] = x_{n-1-i} for the sequence x_k = fn(k % period, k)."""
    return [fn((n - 1 - i) % period, n - 1 - i) for i in range(n)]


def polynomials(rng, period, common_degree, kind_degree):
    """x_k = a common polynomial in t = k / 10 plus one per kind of solve (k % period)."""
    common = rng.normal(size=common_degree + 1)
    kind = rng.normal(size=(period, max(kind_degree, 0) + 1)) if kind_degree >= 0 else None

    def fn(q, k):
        t = 0.1 * k
        v = sum(c * t ** p for p, c in enumerate(common))
        if kind is not None:
            v += sum(c * t ** p for p, c in enumerate(kind[q]))
        return v

    return fn


@pytest.mark.parametrize("period", [1, 2])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_exact_for_polynomials_of_the_order_and_not_beyond(order, period):
    rng = np.random.default_rng(10 * order + period)
    for trial in range(20):
        n = 40 + trial
        fn = polynomials(rng, period, order, order - 1)
        exact = fn(n % period, n)
        got = orc.ws_extrapolate(order, period, history(fn, n, period))
        assert abs(got - exact) <= 1e-9 * max(1.0, abs(exact))
        # one degree more in the common part, or (two kinds) in the per-kind part: not exact
        higher = [polynomials(rng, period, order + 1, order - 1)]
        if period == 2:
            higher.append(polynomials(rng, period, order, order))
        for fn in higher:
            miss = orc.ws_extrapolate(order, period, history(fn, n, period))
            assert abs(miss - fn(n % period, n)) > 1e-7


@pytest.mark.parametrize("order,window", [(2, 4), (3, 6), (3, 8), (4, 10)])
def test_wider_windows_stay_exact_to_their_order(order, window):
    rng = np.random.default_rng(100 * order + window)
    for trial in range(10):
        n = 60 + trial
        fn = polynomials(rng, 2, order, order - 1)
        got = orc.ws_extrapolate(order, 2, history(fn, n, 2), window=window)
        exact = fn(n % 2, n)
        assert abs(got - exact) <= 1e-8 * max(1.0, abs(exact))


def least_norm(order, window):
    """numpy's least-norm solution of V c = e_0, V[r][j] = (-j)^r (r < order, j = 1..window)."""
    V = np.array([[(-j) ** r for j in range(1, window + 1)] for r in range(order)], dtype=float)
    e = np.zeros(order)
    e[0] = 1.0
    return np.linalg.lstsq(V, e, rcond=None)[0]


def test_weights_are_the_least_norm_exact_ones():
    from wavesim_amd import solver as gpu

    have_lib = os.path.exists(gpu.LIB_PATH)
    for order in range(1, 6):
        for window in range(order, 13):
            want = least_norm(order, window)
            got = orc.ws_weights(order, window)
            assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max()), (
                order, window, got, want)
            if window == order:  # the binomial weights, exactly
                binom = [(-1) ** (j + 1) * math.comb(order, j) for j in range(1, order + 1)]
                assert list(got) == binom
            if have_lib:  # the product's host-side weights (no GPU needed), bit for bit:
                # the FP64 path's warm start is held to the oracle's bitwise
                prod = gpu.warm_start_weights(order, window)
                assert np.array_equal(prod, got), (order, window, prod, got)


def test_wider_windows_let_less_noise_through():
    # sum c_j^2 is what carries the past solutions' rounding into the start
    assert np.sum(orc.ws_weights(3, 8) ** 2) < 0.4 * np.sum(orc.ws_weights(3, 3) ** 2)


def test_order_zero_is_the_previous_solution():
    past = [3.25, -1.0, 7.0, 2.0, 0.5, 9.0, 4.0]
    for period in (1, 2):
        assert orc.ws_extrapolate(0, period, past) == 3.25


@pytest.mark.parametrize("order", [1, 2, 3])
def test_error_falls_as_dt_to_the_order_plus_one(order):
    # two interleaved kinds of solve, as the RK2 stages: f(t) at t, and f + dt g (the predictor
    # state's O(dt) offset) at t + dt
    def err(dt, n):
        def fn(q, k):
            t = 1.3 + (k // 2) * dt + q * dt
            return math.sin(0.7 * t) + q * dt * 0.3 * math.cos(0.5 * t)
        return abs(orc.ws_extrapolate(order, 2, history(fn, n, 2)) - fn(n % 2, n))

    for n in (20, 21):  # the next solve of either kind
        rate = math.log2(err(0.005, n) / err(0.0025, n))
        assert abs(rate - (order + 1)) < 0.1, (n, rate)


def test_missing_history_reads_as_zero_steps():
    # before the record fills the start falls back to lower orders, never to garbage
    assert orc.ws_extrapolate(3, 2, [5.0]) == 5.0
    assert orc.ws_extrapolate(3, 2, [5.0, 4.0, 2.0]) == 5.0 + 3.0 * 2.0


def test_a_run_keeps_its_waves_in_fewer_iterations():
    w = configs.c1_rough_bed(nx=128, ny=48, steps=60)
    runs = {}
    for order in (0, 3):
        o = w.solver(lambda g, b: orc.OracleSolver(g, b)).with_warm_start(order)
        s = w.still_water()
        s.time = 40.0  # mid-ramp: the maker drives the dispersion
        for _ in range(w.steps):
            o.step(s, o.stable_dt(s))
        st = o.dispersion_stats()
        runs[order] = (s, st.iterations / st.solves, st.unconverged)
    (s0, it0, u0), (s3, it3, u3) = runs[0], runs[3]
    assert u0 == 0 and u3 == 0
    assert it3 < 0.7 * it0, (it0, it3)
    assert np.max(np.abs(s3.arrays()[0] - s0.arrays()[0])) < 1e-6
