"""Pins the network oracle (oracle/unet_oracle.py) before it checks the GPU: each operation
against what the mathematics fixes, so that a dropped term, a flipped tap or a transposed
operand fails here."""

from __future__ import annotations

import numpy as np

from oracle import unet_oracle as uo


def _rng(seed=0):
    return np.random.default_rng(seed)


def test_conv_matches_explicit_loops():
    r = _rng(1)
    H, W, ci, co = 5, 6, 3, 2
    x = r.normal(size=(H, W, ci))
    w = r.normal(size=(9, ci, co))
    b = r.normal(size=co)
    got = uo.conv(x, w, b)
    want = np.zeros((H, W, co))
    for y in range(H):
        for xx in range(W):
            for o in range(co):
                s = b[o]
                for t in range(9):
                    yy, xs = y + t // 3 - 1, xx + t % 3 - 1
                    if 0 <= yy < H and 0 <= xs < W:
                        s += sum(w[t, i, o] * x[yy, xs, i] for i in range(ci))
                want[y, xx, o] = s
    assert np.max(np.abs(got - want)) < 1e-12


def test_transposed_conv_and_weight_gradient_are_the_adjoints():
    r = _rng(2)
    for taps in (1, 9):
        x = r.normal(size=(8, 8, 4))
        w = r.normal(size=(taps, 4, 3))
        y = r.normal(size=(8, 8, 3))
        lhs = np.sum(uo.conv(x, w) * y)
        assert abs(lhs - np.sum(x * uo.conv_t(y, w))) < 1e-10 * abs(lhs)
        assert abs(lhs - np.sum(w * uo.wgrad(x, y, taps))) < 1e-10 * abs(lhs)


def test_pool_and_upsample_adjoints():
    r = _rng(3)
    x = r.normal(size=(8, 6, 2))
    p = r.normal(size=(4, 3, 2))
    assert abs(np.sum(uo.pool(x) * p) - np.sum(x * uo.pool_t(p))) < 1e-12
    assert abs(np.sum(uo.up(p) * x) - np.sum(p * uo.up_t(x))) < 1e-12
    assert np.allclose(uo.pool(uo.up(p)), p)


def test_bf16_rounds_to_nearest_even():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to the even mantissa (1)
    assert uo.bf16(1.0 + 2.0 ** -8) == 1.0
    assert uo.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert uo.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7


def test_backward_matches_finite_differences():
    r = _rng(4)
    H, W = 16, 16
    params = np.concatenate([
        np.concatenate([r.normal(0.0, 1.0 / np.sqrt(t * ci), t * ci * co), r.normal(0, 0.1, co)])
        for _, t, ci, co, _ in uo.LAYERS])
    x0 = r.normal(size=(H, W, 8))
    target = r.normal(size=(H, W, 2))
    weight = (r.random((H, W)) > 0.2).astype(float)

    def L(p):
        return uo.loss(uo.forward(p, x0)["out"], target, weight)[0]

    a = uo.forward(params, x0)
    _, dout = uo.loss(a["out"], target, weight)
    pg, _ = uo.backward(params, a, dout)
    g = uo.flat_grads(pg)
    assert g.size == params.size
    o, eps = 0, 1e-6
    for name, taps, ci, co, _ in uo.LAYERS:
        n = taps * ci * co + co
        for k in [o, o + n // 2, o + n - 1]:  # first weight, a middle one, the last bias
            p1, p2 = params.copy(), params.copy()
            p1[k] += eps
            p2[k] -= eps
            num = (L(p1) - L(p2)) / (2 * eps)
            assert abs(num - g[k]) <= 1e-5 * max(1.0, abs(num)), (name, k, num, g[k])
        o += n
