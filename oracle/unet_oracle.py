"""TEST INFRASTRUCTURE ONLY: a numpy FP64 restatement of the learned preconditioner's network.

The network has no counterpart in the reference (include/wavecore_unet.h describes it); this
file is the checker of the HIP kernels in wavesim_amd/csrc/wc_unet.hip: every operation of the
forward and backward passes, on unpadded (H, W, C) arrays with zero padding outside, in FP64.
``bf16`` rounds like the GPU (round to nearest even from FP32). Only tests/ load it.

Its own pins (tests/test_unet_oracle.py): the convolution against explicit loops, the
transposed convolution and the weight gradient against the adjoint identities they must
satisfy, pooling and upsampling against theirs, and the whole backward pass against central
finite differences of the loss.
"""

from __future__ import annotations

import numpy as np

# (name, taps, cin_real, cout, relu) in parameter order; matches wc_unet.hip's table.
LAYERS = [("E0", 9, 5, 16, True), ("E1", 9, 16, 32, True), ("E2", 9, 32, 64, True),
          ("E3", 9, 64, 128, True), ("B", 9, 128, 128, True), ("T3", 1, 128, 64, False),
          ("D2", 9, 64, 64, True), ("T2", 1, 64, 32, False), ("D1", 9, 32, 32, True),
          ("T1", 1, 32, 16, False), ("D0", 9, 16, 16, True), ("O", 1, 16, 2, False)]


def bf16(a) -> np.ndarray:
    """Round to bfloat16 (nearest even, through FP32), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + (((u >> 16) & 1) + 0x7FFF)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray) -> dict:
    """{name: (W[taps, cin, cout], b[cout])} from the flat FP32 parameter vector."""
    out, o = {}, 0
    for name, taps, cin, cout, _ in LAYERS:
        n = taps * cin * cout
        w = np.asarray(params[o:o + n], dtype=np.float64).reshape(taps, cin, cout)
        b = np.asarray(params[o + n:o + n + cout], dtype=np.float64)
        out[name] = (w, b)
        o += n + cout
    return out


def param_count() -> int:
    return sum(t * ci * co + co for _, t, ci, co, _ in LAYERS)


def _offsets(taps):
    return [(t // 3 - 1, t % 3 - 1) for t in range(9)] if taps == 9 else [(0, 0)]


def _shift(x, dy, dx):
    """s[y, x] = x[y + dy, x + dx], zero outside."""
    H, W = x.shape[:2]
    p = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    return p[1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]


def conv(x, w, b=None):
    """out[y, x, o] = b[o] + sum_{t, i} w[t, i, o] x[y + dy_t, x + dx_t, i]."""
    taps = w.shape[0]
    out = np.zeros(x.shape[:2] + (w.shape[2],))
    for t, (dy, dx) in enumerate(_offsets(taps)):
        out += _shift(x, dy, dx) @ w[t]
    return out if b is None else out + b


def conv_t(dy_, w):
    """The adjoint of conv in its input: dx[y, x, i] = sum_{t, o} w[t, i, o] dy[y - dy_t, x - dx_t, o]."""
    taps = w.shape[0]
    out = np.zeros(dy_.shape[:2] + (w.shape[1],))
    for t, (dy, dx) in enumerate(_offsets(taps)):
        out += _shift(dy_, -dy, -dx) @ w[t].T
    return out


def wgrad(x, dy_, taps):
    """d <conv(x, w), dy> / d w[t, i, o] = sum_p x[p + d_t, i] dy[p, o]."""
    cin, cout = x.shape[2], dy_.shape[2]
    out = np.zeros((taps, cin, cout))
    for t, (dy, dx) in enumerate(_offsets(taps)):
        out[t] = _shift(x, dy, dx).reshape(-1, cin).T @ dy_.reshape(-1, cout)
    return out


def pool(x):
    return 0.25 * ((x[0::2, 0::2] + x[0::2, 1::2]) + (x[1::2, 0::2] + x[1::2, 1::2]))


def pool_t(dp):
    return 0.25 * up(dp)


def up(t):
    return np.repeat(np.repeat(t, 2, axis=0), 2, axis=1)


def up_t(du):
    return (du[0::2, 0::2] + du[0::2, 1::2]) + (du[1::2, 0::2] + du[1::2, 1::2])


def relu(x):
    return np.maximum(x, 0.0)


def forward(params, x0, rnd=None) -> dict:
    """The network's activations from the input x0 (H, W, 5 or 8), each passed through rnd
    (e.g. bf16) as the GPU stores it; 'out' is the FP32-free output (H, W, 2)."""
    r = rnd or (lambda a: a)
    P = unpack(params)
    a = {"x0": r(x0[..., :5])}

    def layer(name, x):
        w, b = P[name]
        return conv(x, w, b)

    a["e0"] = r(relu(layer("E0", a["x0"])))
    a["p1"] = r(pool(a["e0"]))
    a["e1"] = r(relu(layer("E1", a["p1"])))
    a["p2"] = r(pool(a["e1"]))
    a["e2"] = r(relu(layer("E2", a["p2"])))
    a["p3"] = r(pool(a["e2"]))
    a["e3"] = r(relu(layer("E3", a["p3"])))
    a["y3"] = r(relu(layer("B", a["e3"])))
    a["t3"] = r(layer("T3", a["y3"]))
    a["u2"] = r(up(a["t3"]) + a["e2"])
    a["y2"] = r(relu(layer("D2", a["u2"])))
    a["t2"] = r(layer("T2", a["y2"]))
    a["u1"] = r(up(a["t2"]) + a["e1"])
    a["y1"] = r(relu(layer("D1", a["u1"])))
    a["t1"] = r(layer("T1", a["y1"]))
    a["u0"] = r(up(a["t1"]) + a["e0"])
    a["y0"] = r(relu(layer("D0", a["u0"])))
    a["out"] = layer("O", a["y0"])
    return a


def loss(out, target, weight):
    """sum(weight |out - target|^2) / sum(weight) over cells; and d loss / d out."""
    d = out - target
    sw = weight.sum()
    val = float((weight[..., None] * d * d).sum() / sw)
    return val, 2.0 * weight[..., None] * d / sw


def backward(params, a, dout, rnd=None) -> tuple[dict, dict]:
    """Gradients from d loss / d out through the activations a (as forward returned them):
    ({name: (dW, db)}, {tensor: gradient}); gradients of ReLU layers are taken with respect to
    their pre-activation, each passed through rnd as the GPU stores it."""
    r = rnd or (lambda x: x)
    P = unpack(params)
    g, pg = {}, {}

    def wb(name, x, dy_):
        pg[name] = (wgrad(x, dy_, P[name][0].shape[0]), dy_.reshape(-1, dy_.shape[2]).sum(0))

    wb("O", a["y0"], dout)
    g["y0"] = r(conv_t(dout, P["O"][0]) * (a["y0"] > 0))
    wb("D0", a["u0"], g["y0"])
    g["u0"] = r(conv_t(g["y0"], P["D0"][0]))
    g["t1"] = r(up_t(g["u0"]))
    wb("T1", a["y1"], g["t1"])
    g["y1"] = r(conv_t(g["t1"], P["T1"][0]) * (a["y1"] > 0))
    wb("D1", a["u1"], g["y1"])
    g["u1"] = r(conv_t(g["y1"], P["D1"][0]))
    g["t2"] = r(up_t(g["u1"]))
    wb("T2", a["y2"], g["t2"])
    g["y2"] = r(conv_t(g["t2"], P["T2"][0]) * (a["y2"] > 0))
    wb("D2", a["u2"], g["y2"])
    g["u2"] = r(conv_t(g["y2"], P["D2"][0]))
    g["t3"] = r(up_t(g["u2"]))
    wb("T3", a["y3"], g["t3"])
    g["y3"] = r(conv_t(g["t3"], P["T3"][0]) * (a["y3"] > 0))
    wb("B", a["e3"], g["y3"])
    g["e3"] = r(conv_t(g["y3"], P["B"][0]) * (a["e3"] > 0))
    wb("E3", a["p3"], g["e3"])
    g["p3"] = r(conv_t(g["e3"], P["E3"][0]))
    g["e2"] = r((g["u2"] + pool_t(g["p3"])) * (a["e2"] > 0))
    wb("E2", a["p2"], g["e2"])
    g["p2"] = r(conv_t(g["e2"], P["E2"][0]))
    g["e1"] = r((g["u1"] + pool_t(g["p2"])) * (a["e1"] > 0))
    wb("E1", a["p1"], g["e1"])
    g["p1"] = r(conv_t(g["e1"], P["E1"][0]))
    g["e0"] = r((g["u0"] + pool_t(g["p1"])) * (a["e0"] > 0))
    wb("E0", a["x0"], g["e0"])
    return pg, g


def flat_grads(pg: dict) -> np.ndarray:
    """The parameter gradients in the flat parameter order."""
    return np.concatenate([np.concatenate([pg[n][0].ravel(), pg[n][1]]) for n, *_ in LAYERS])
