"""CPU checks of the boundary: the C-ABI library loads and exports every entry point that
include/wavecore_hip.h declares (no compute: there is no GPU here), and the Python mirror
describes each of them."""

from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from wavesim_amd import solver as gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wavecore_hip.h")


def declared() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared()
    for must in ("wc_solver_new", "wc_step", "wc_stable_dt", "wc_limiting_cell",
                 "wc_breaking_indicator", "wc_dispersion_stats", "wc_state_set", "wc_state_get"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(gpu.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (wc_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    L = ctypes.CDLL(gpu.LIB_PATH)
    for n in declared():
        assert hasattr(L, n)


def test_python_mirror_binds_every_declared_symbol():
    assert sorted(gpu.SIGNATURES) == declared()


@pytest.mark.skipif(not os.path.exists(gpu.LIB_PATH), reason="library not built")
def test_free_functions_work_without_a_gpu():
    # Host-only entry points of the ABI.
    lw = gpu.linear_wave(2 * 3.141592653589793 / 8.0, 5.0, True)
    assert 0.0 < lw.group_speed < lw.phase_speed
    with pytest.raises(ValueError):
        gpu.linear_wave(2 * 3.141592653589793 / 1.0, 5.0, True)
    assert gpu.manning_factor(1.5, 0.8, 0.2, 0.0) == 1.0


UNET_HEADER = os.path.join(ROOT, "include", "wavecore_unet.h")


def declared_unet() -> list[str]:
    text = open(UNET_HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wcu_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(gpu.LIB_PATH), reason="library not built")
def test_library_exports_the_network_api():
    from wavesim_amd import unet

    out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (wcu_[a-z0-9_]+)$", out, flags=re.M))
    names = declared_unet()
    assert "wcu_net_forward" in names and "wcu_net_backward" in names
    assert [n for n in names if n not in exported] == []
    assert sorted(unet.SIGNATURES) == names
    # host-only entry points: the parameter layout
    assert unet.param_count() == sum(d["weights"] + d["biases"] for d in unet.layers())
    from oracle import unet_oracle as uo

    assert unet.param_count() == uo.param_count()
    for d, (name, taps, ci, co, _) in zip(unet.layers(), uo.LAYERS):
        assert (d["name"], d["taps"], d["cin"], d["cout"]) == (name, taps, ci, co)


RUN_HEADER = os.path.join(ROOT, "include", "wavesim_run.h")


def test_library_exports_the_run_writer():
    from wavesim_amd import run

    text = re.sub(r"/\*.*?\*/", "", open(RUN_HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(wc_[a-z0-9_]+)\s*\(", text)))
    assert "wc_run_create" in names and "wc_run_frame_from" in names
    assert sorted(run.SIGNATURES) == names
    if os.path.exists(gpu.LIB_PATH):
        out = subprocess.run(["nm", "-D", "--defined-only", gpu.LIB_PATH], capture_output=True,
                             text=True, check=True).stdout
        exported = set(re.findall(r"\bT (wc_[a-z0-9_]+)$", out, flags=re.M))
        assert [n for n in names if n not in exported] == []
