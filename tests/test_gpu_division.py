"""The flux kernels' division by a divisor every cell shares, against the device's own division.

The fifth-order formula divides by 60 and the rates by dx and dy; the flux kernels compute each
divisor's refined reciprocal once and take each quotient in three instructions where the
division's scaling and fix-up steps would change nothing (wc_device.h: quot_of, div_q). The
FP64 path's bit-for-bit parity with the oracle rests on these quotients being the division's:
here both ways on the device, over numerators spread across the whole double range and the
special values, for the divisors the workloads use and a few others; and both against numpy's
IEEE division.
"""

from __future__ import annotations

import numpy as np
import pytest

from wavesim_amd.solver import check_shared_division

pytestmark = pytest.mark.gpu


def numerators(seed: int, n: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    # random bit patterns cover every exponent (subnormals and non-finite values included)
    bits = rng.integers(0, 2**63, size=n, dtype=np.int64).astype(np.uint64)
    bits |= rng.integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63)
    spread = bits.view(np.float64)
    # magnitudes the kernels meet (fluxes, rates, face values), and the fast range's edges
    typical = rng.standard_normal(n) * 10.0 ** rng.uniform(-12, 4, size=n)
    edges = []
    for e in (-900, 700):
        b = 2.0 ** e
        edges += [b, -b, np.nextafter(b, 0.0), np.nextafter(b, np.inf), -np.nextafter(b, 0.0)]
    special = [0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -5e-324, 2.2250738585072014e-308,
               1.7976931348623157e308, -1.7976931348623157e308, 60.0, 1.0, -1.0]
    return np.concatenate([spread, typical, np.array(edges + special)])


@pytest.mark.parametrize("y", [60.0, 3.0, 2.0, 1.5, 0.7, 7.3, 1e-5, 1e5, 2.0 ** -70, -3.0])
def test_shared_reciprocal_gives_the_division_bits(gpu, y):
    x = numerators(int(abs(y) * 1000) % 2**31, 200_000)
    shared, plain = check_shared_division(y, x)
    assert np.array_equal(shared.view(np.uint64), plain.view(np.uint64)), (
        f"{int(np.sum(shared.view(np.uint64) != plain.view(np.uint64)))} quotients differ")
    with np.errstate(all="ignore"):
        ieee = x / y
    # the device's division is IEEE's (subnormal quotients aside, which no kernel here meets)
    normal = np.isfinite(ieee) & ((np.abs(ieee) >= np.finfo(np.float64).tiny) | (ieee == 0.0))
    assert np.array_equal(plain[normal].view(np.uint64), ieee[normal].view(np.uint64))
    assert np.array_equal(np.isnan(plain), np.isnan(ieee))
