"""Builds libwavecore_hip.so in-tree with hipcc for gfx950.

    python -m wavesim_amd.build

Flags: -ffp-contract=off (the FP64 path reproduces the CPU oracle bit for bit, and Rust never
fuses a*b+c), no fast-math (FP64 division and square root stay correctly rounded), and packed
FP32 instructions disabled (they are not used by hand anywhere).
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwavecore_hip.so")
SOURCES = ["wc_api.hip", "wc_stats_host.hip", "wc_nested_host.hip", "wc_batch_host.hip", "wc_kernels.hip", "wc_flux.hip", "wc_fp32.hip", "wc_pcg.hip",
           "wc_pcg1.hip", "wc_pcg2.hip", "wc_relax.hip", "wc_unet.hip", "wc_fpcg.hip",
           "wc_stats.hip", "wc_run.hip", "wc_eddy.hip", "wc_noalloc.hip", "wc_stream.hip"]
HEADERS = ["wc_host.h", "wc_device.h", "wc_json.h", "wc_kernels.h", "wc_math.h", "wc_run.h", "wc_stage.h",
           "wc_tile16.h"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # The host pass does not know the packed-FP32 feature flag and says so; nothing else is
    # expected on stderr.
    noise = [ln for ln in res.stderr.splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "warning generated" not in ln]
    return res, noise


def build(verbose: bool = False) -> str:
    """Each source compiled to an object on its own (in parallel), then one link."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = [f for f in FLAGS if f != "-shared"]
    jobs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        jobs.append((obj, [hipcc(), *flags, "-c", "-o", obj, os.path.join(CSRC, src)]))
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(lambda j: _run(j[1], verbose), jobs))
    for (obj, _), (res, noise) in zip(jobs, results):
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed compiling {os.path.basename(obj)}")
        if noise and verbose:
            sys.stderr.write("\n".join(noise) + "\n")
    res, noise = _run([hipcc(), *FLAGS, "-o", LIB + ".tmp", *[o for o, _ in jobs]], verbose)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libwavecore_hip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
