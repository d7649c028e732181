"""The learned preconditioner's network (include/wavecore_unet.h) from Python.

The network runs in the library's HIP kernels (BF16 MFMA convolutions, FP32 accumulation);
torch provides the device memory and the stream, nothing else. ``UNet`` wraps one network
instance: load parameters, build the input from grid-layout arrays, forward, loss, backward;
``Adam`` updates FP32 parameters with the library's kernel.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .solver import _check, lib

_U64 = ctypes.c_uint64
_I32 = ctypes.c_int32
_F = ctypes.c_float
_VP = ctypes.c_void_p

# Every entry point of include/wavecore_unet.h with its ctypes signature.
SIGNATURES = {
    "wcu_param_count": (_U64, []),
    "wcu_layer": (_I32, [_I32, ctypes.POINTER(_U64), ctypes.POINTER(_U64), ctypes.POINTER(_U64),
                         ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "wcu_net_new": (_I32, [_U64, _U64, ctypes.POINTER(_VP)]),
    "wcu_net_free": (None, [_VP]),
    "wcu_net_set_params": (_I32, [_VP, _VP, _VP]),
    "wcu_net_input": (_I32, [_VP, _U64, _U64, _U64, _U64, _VP, _VP, _VP, _VP, _VP, _F, _F,
                             _VP]),
    "wcu_net_forward": (_I32, [_VP, _VP, _VP]),
    "wcu_net_loss": (_I32, [_VP, _VP, _VP, ctypes.POINTER(ctypes.c_double), _VP]),
    "wcu_net_backward": (_I32, [_VP, _VP, _VP]),
    "wcu_adam": (_I32, [_U64, _VP, _VP, _VP, _VP, _F, _F, _F, _F, _I32, _VP]),
    "wcu_net_tensor": (_I32, [_VP, _I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_U64),
                              ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
}

LAYER_NAMES = ["E0", "E1", "E2", "E3", "B", "T3", "D2", "T2", "D1", "T1", "D0", "O"]
TENSOR_NAMES = ["x0", "e0", "p1", "e1", "p2", "e2", "p3", "e3", "y3", "t3", "u2", "y2", "t2",
                "u1", "y1", "t1", "u0", "y0"]
GRAD = 100  # tensor id offset of gradients

_bound = False


def ulib() -> ctypes.CDLL:
    global _bound
    L = lib()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


def _ptr(t) -> int:
    return int(t.data_ptr())


def _stream() -> int:
    import torch

    return int(torch.cuda.current_stream().cuda_stream)


def param_count() -> int:
    return int(ulib().wcu_param_count())


def layers() -> list[dict]:
    """Per layer: name, parameter offset, weight and bias counts, taps, in and out channels."""
    out = []
    for l, name in enumerate(LAYER_NAMES):
        o, w, b = _U64(), _U64(), _U64()
        t, ci, co = _I32(), _I32(), _I32()
        _check(ulib().wcu_layer(l, ctypes.byref(o), ctypes.byref(w), ctypes.byref(b),
                                ctypes.byref(t), ctypes.byref(ci), ctypes.byref(co)))
        out.append({"name": name, "offset": o.value, "weights": w.value, "biases": b.value,
                    "taps": t.value, "cin": ci.value, "cout": co.value})
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (std sqrt(2 / fan_in)), zero biases; the output layer starts small."""
    rng = np.random.default_rng(seed)
    p = np.zeros(param_count(), dtype=np.float32)
    for d in layers():
        fan_in = d["taps"] * d["cin"]
        std = np.sqrt(2.0 / fan_in) if d["name"] != "O" else 0.1 / np.sqrt(fan_in)
        if d["name"] in ("T3", "T2", "T1"):
            std = np.sqrt(1.0 / fan_in)
        p[d["offset"]:d["offset"] + d["weights"]] = rng.normal(0.0, std, d["weights"])
    return p


class UNet:
    """One network instance for grids of up to nx x ny cells (include/wavecore_unet.h)."""

    def __init__(self, nx: int, ny: int):
        import torch

        self.nx, self.ny = nx, ny
        h = _VP()
        _check(ulib().wcu_net_new(nx, ny, ctypes.byref(h)))
        self._h = h
        self.out = torch.zeros(2 * nx * ny, dtype=torch.float32, device="cuda")

    def __del__(self):
        if getattr(self, "_h", None):
            ulib().wcu_net_free(self._h)
            self._h = None

    def set_params(self, params) -> None:
        """params: a CUDA float32 tensor of param_count() values."""
        assert params.dtype.is_floating_point and params.numel() == param_count()
        _check(ulib().wcu_net_set_params(self._h, _ptr(params), _stream()))

    def input(self, h, bgx, bgy, rx, ry, pitch: int, ghost: int, h_scale: float,
              r_scale: float, nx: int | None = None, ny: int | None = None) -> None:
        """Input channels from grid-layout CUDA float32 tensors (row pitch ``pitch``, interior
        at (ghost, ghost))."""
        _check(ulib().wcu_net_input(self._h, nx or self.nx, ny or self.ny, pitch, ghost,
                                    _ptr(h), _ptr(bgx), _ptr(bgy), _ptr(rx), _ptr(ry),
                                    h_scale, r_scale, _stream()))

    def forward(self):
        _check(ulib().wcu_net_forward(self._h, _ptr(self.out), _stream()))
        return self.out

    def loss(self, target, weight) -> float:
        out = ctypes.c_double()
        _check(ulib().wcu_net_loss(self._h, _ptr(target), _ptr(weight), ctypes.byref(out),
                                   _stream()))
        return out.value

    def backward(self, grads) -> None:
        _check(ulib().wcu_net_backward(self._h, _ptr(grads), _stream()))

    def tensor(self, tid: int) -> np.ndarray:
        """An internal tensor (see include/wavecore_unet.h), interior only, as (H, W, C)."""
        h, w, c = _U64(), _U64(), _U64()
        _check(ulib().wcu_net_tensor(self._h, tid, None, ctypes.byref(h), ctypes.byref(w),
                                     ctypes.byref(c)))
        out = np.empty(h.value * w.value * c.value, dtype=np.float32)
        _check(ulib().wcu_net_tensor(self._h, tid,
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None,
                                     None, None))
        return out.reshape(h.value, w.value, c.value)


class Adam:
    """Adam (Kingma and Ba, 2015) on a CUDA float32 parameter tensor, in the library's kernel."""

    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8):
        import torch

        self.params = params
        self.m = torch.zeros_like(params)
        self.v = torch.zeros_like(params)
        self.lr, self.betas, self.eps, self.t = lr, betas, eps, 0

    def step(self, grads) -> None:
        self.t += 1
        _check(ulib().wcu_adam(self.params.numel(), _ptr(self.params), _ptr(grads), _ptr(self.m),
                               _ptr(self.v), self.lr, self.betas[0], self.betas[1], self.eps,
                               self.t, _stream()))
