"""The network's HIP kernels (wavesim_amd/csrc/wc_unet.hip) against the FP64 oracle
(oracle/unet_oracle.py), layer by layer: every operation is recomputed on the CPU from the
GPU's own BF16 inputs, so BF16 rounding does not accumulate through the check.

Bars: a BF16 output (activations, gradients of activations) equals the oracle's value rounded
to BF16 on >= 99% of elements and is within 2^-7 of it relatively everywhere (one BF16 ulp,
the FP32 accumulation order decides ties); FP32 weight gradients are within 1e-4 of the
oracle in relative Frobenius norm. Training lowers the loss.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import unet_oracle as uo

pytestmark = pytest.mark.gpu

NX, NY, GH = 256, 64, 3


def setup_net(seed=0):
    import torch

    from wavesim_amd.unet import UNet, init_params

    net = UNet(NX, NY)
    p = init_params(seed)
    params = torch.tensor(p, device="cuda")
    net.set_params(params)
    r = np.random.default_rng(seed + 10)
    pitch = NX + 2 * GH
    grid = [r.normal(size=(NY + 2 * GH) * pitch).astype(np.float32) for _ in range(5)]
    grid[0] = np.abs(grid[0]) * 10.0  # depth
    dev = [torch.tensor(a, device="cuda") for a in grid]
    net.input(*dev, pitch=pitch, ghost=GH, h_scale=0.1, r_scale=1.0)
    return net, p, grid


def bf16_close(gpu, ref, label):
    want = uo.bf16(ref)
    exact = np.mean(gpu == want)
    scale = np.maximum(np.abs(ref), 1e-3 * max(np.max(np.abs(ref)), 1e-30))
    rel = np.max(np.abs(gpu - ref) / scale)
    print(f"{label}: exact {exact:.5f}, max rel {rel:.2e}", flush=True)
    assert exact >= 0.99, (label, exact)
    assert rel <= 2.0 ** -7, (label, rel)


def test_forward_matches_the_oracle_layer_by_layer():
    import torch

    from wavesim_amd.unet import TENSOR_NAMES

    net, p, grid = setup_net(1)
    out = net.forward()
    torch.cuda.synchronize()
    T = {n: net.tensor(i) for i, n in enumerate(TENSOR_NAMES)}
    # the input channels
    pitch = NX + 2 * GH
    inner = [a.reshape(NY + 2 * GH, pitch)[GH:GH + NY, GH:GH + NX] for a in grid]
    x0 = np.stack([inner[0] * np.float32(0.1), inner[1], inner[2], inner[3], inner[4]], -1)
    bf16_close(T["x0"][..., :5], x0.astype(np.float64), "x0")
    assert np.all(T["x0"][..., 5:] == 0)
    P = uo.unpack(p)

    def lay(name, x):
        w, b = P[name]
        return uo.conv(x, uo.bf16(w), b)

    checks = [
        ("e0", uo.relu(lay("E0", T["x0"][..., :5]))), ("p1", uo.pool(T["e0"])),
        ("e1", uo.relu(lay("E1", T["p1"]))), ("p2", uo.pool(T["e1"])),
        ("e2", uo.relu(lay("E2", T["p2"]))), ("p3", uo.pool(T["e2"])),
        ("e3", uo.relu(lay("E3", T["p3"]))), ("y3", uo.relu(lay("B", T["e3"]))),
        ("t3", lay("T3", T["y3"])), ("u2", uo.up(T["t3"]) + T["e2"]),
        ("y2", uo.relu(lay("D2", T["u2"]))), ("t2", lay("T2", T["y2"])),
        ("u1", uo.up(T["t2"]) + T["e1"]), ("y1", uo.relu(lay("D1", T["u1"]))),
        ("t1", lay("T1", T["y1"])), ("u0", uo.up(T["t1"]) + T["e0"]),
        ("y0", uo.relu(lay("D0", T["u0"]))),
    ]
    for name, ref in checks:
        bf16_close(T[name].astype(np.float64), ref, name)
    w, b = P["O"]
    ref_out = uo.conv(T["y0"].astype(np.float64), w, b)
    got = out.cpu().numpy().reshape(NY, NX, 2)
    assert np.max(np.abs(got - ref_out)) <= 1e-5 * max(1.0, np.max(np.abs(ref_out)))


def test_backward_matches_the_oracle_layer_by_layer():
    import torch

    from wavesim_amd.unet import GRAD, TENSOR_NAMES, param_count

    net, p, _ = setup_net(2)
    out = net.forward()
    r = np.random.default_rng(5)
    target = r.normal(size=2 * NX * NY).astype(np.float32)
    weight = (r.random(NX * NY) > 0.1).astype(np.float32)
    net.loss(torch.tensor(target, device="cuda"), torch.tensor(weight, device="cuda"))
    grads = torch.zeros(param_count(), device="cuda")
    net.backward(grads)
    torch.cuda.synchronize()
    T = {n: net.tensor(i).astype(np.float64) for i, n in enumerate(TENSOR_NAMES)}
    G = {n: net.tensor(GRAD + i).astype(np.float64) for i, n in enumerate(TENSOR_NAMES) if i}
    o = out.cpu().numpy().astype(np.float64).reshape(NY, NX, 2)
    _, dout = uo.loss(o, target.reshape(NY, NX, 2).astype(np.float64),
                      weight.reshape(NY, NX).astype(np.float64))
    P = uo.unpack(p)
    W = {k: uo.bf16(v[0]) for k, v in P.items()}
    checks = [
        ("y0", uo.conv_t(dout, P["O"][0]) * (T["y0"] > 0)),
        ("u0", uo.conv_t(G["y0"], W["D0"])), ("t1", uo.up_t(G["u0"])),
        ("y1", uo.conv_t(G["t1"], W["T1"]) * (T["y1"] > 0)),
        ("u1", uo.conv_t(G["y1"], W["D1"])), ("t2", uo.up_t(G["u1"])),
        ("y2", uo.conv_t(G["t2"], W["T2"]) * (T["y2"] > 0)),
        ("u2", uo.conv_t(G["y2"], W["D2"])), ("t3", uo.up_t(G["u2"])),
        ("y3", uo.conv_t(G["t3"], W["T3"]) * (T["y3"] > 0)),
        ("e3", uo.conv_t(G["y3"], W["B"]) * (T["e3"] > 0)),
        ("p3", uo.conv_t(G["e3"], W["E3"])),
        ("e2", (G["u2"] + uo.pool_t(G["p3"])) * (T["e2"] > 0)),
        ("p2", uo.conv_t(G["e2"], W["E2"])),
        ("e1", (G["u1"] + uo.pool_t(G["p2"])) * (T["e1"] > 0)),
        ("p1", uo.conv_t(G["e1"], W["E1"])),
        ("e0", (G["u0"] + uo.pool_t(G["p1"])) * (T["e0"] > 0)),
    ]
    for name, ref in checks:
        bf16_close(G[name], ref, "d" + name)
    # parameter gradients from the GPU's own tensors
    pg = {
        "O": (T["y0"], dout), "D0": (T["u0"], G["y0"]), "T1": (T["y1"], G["t1"]),
        "D1": (T["u1"], G["y1"]), "T2": (T["y2"], G["t2"]), "D2": (T["u2"], G["y2"]),
        "T3": (T["y3"], G["t3"]), "B": (T["e3"], G["y3"]), "E3": (T["p3"], G["e3"]),
        "E2": (T["p2"], G["e2"]), "E1": (T["p1"], G["e1"]), "E0": (T["x0"][..., :5], G["e0"]),
    }
    g = grads.cpu().numpy().astype(np.float64)
    off = 0
    for name, taps, ci, co, _ in uo.LAYERS:
        x, dy = pg[name]
        dw = uo.wgrad(x, dy, taps)
        db = dy.reshape(-1, co).sum(0)
        n = taps * ci * co
        for lab, got, ref in (("W", g[off:off + n], dw.ravel()), ("b", g[off + n:off + n + co], db)):
            err = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
            print(f"d{lab} {name}: rel {err:.2e}", flush=True)
            assert err <= 1e-4, (name, lab, err)
        off += n + co


def test_training_lowers_the_loss():
    import torch

    from wavesim_amd.unet import Adam, param_count

    net, p, grid = setup_net(3)
    params = torch.tensor(p, device="cuda")
    net.set_params(params)
    pitch = NX + 2 * GH
    inner = [a.reshape(NY + 2 * GH, pitch)[GH:GH + NY, GH:GH + NX] for a in grid]
    # a smooth linear map of the residual channels as the target
    target = np.stack([0.5 * inner[3] - 0.2 * inner[4], 0.3 * inner[4]], -1).astype(np.float32)
    tgt = torch.tensor(target.ravel(), device="cuda")
    wt = torch.ones(NX * NY, device="cuda")
    grads = torch.zeros(param_count(), device="cuda")
    opt = Adam(params, lr=2e-3)
    losses = []
    for _ in range(60):
        net.forward()
        losses.append(net.loss(tgt, wt))
        net.backward(grads)
        opt.step(grads)
        net.set_params(params)
    print("losses", losses[0], losses[-1], flush=True)
    assert np.isfinite(losses[-1])
    assert losses[-1] < 0.3 * losses[0]
