"""Throughput of wavecore's time step on one MI355X: cell-updates/s, PCG iterations per solve,
and the roofline of the dominant kernel, with the CPU FP64 path timed beside it.

    python bench.py [--steps K] [--warmup W] [--spinup S] [--workload rough_bed] [--gpus 1]

Workload (default): the dispersive fifth-order run, 1024x512 cells at 3 m (1:40 plane from
20 m plus 16 random Fourier modes, seed 1), SGN by PCG to 1e-8, breaking switch on, FP64.
A step is one full Solver::step (both RK stages, every PCG iteration) with dt = stable_dt,
driven on the device by wc_run_steps. The timed window starts from a mid-run state: S untimed
spin-up steps (the wave has crossed the shelf by then), W untimed warm-up steps, then K timed.
Inputs are resident in HBM for the whole window.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s (FV+SGN step) at equal accuracy; PCG iters/solve; HBM GB/s"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per interior cell of one launch of each kernel family (FP64: each
# array read or written once, halo re-reads excluded), for the roofline.
BYTES_PER_CELL = {
    "cg_matvec": 8 * (2 + 1 + 1 + 2),      # read p(2), c, h; write Ap(2)
    "cg_update": 8 * (2 + 2 + 2 + 2 + 2 + 2 + 2),  # read w, p, r, Ap, diag; write w, r
    "cg_pupdate": 8 * (2 + 2 + 2 + 2),     # read r, diag, p; write p
    "cg_init": 8 * (2 + 1 + 1 + 2 + 2 + 2 + 2),  # read w, c, h, b, diag; write r, p
    "sgn_setup": 8 * (7 + 5),              # read h, hu, hv, phi, lh, lhu, lhv; write c, diag, b
    "flux": 8 * (5 + 3),                   # read h, hu, hv, b, phi; write lh, lhu, lhv
    "sgn_finish": 8 * (3 + 3 + 2 + 3 + 3 + 1) // 2 * 2,  # stage in + rates + w + (s, b) + out
    "switch": 8 * 3,                       # read h, b; write phi (per padded cell ~ interior)
    "fill_state": 0,
}


def parse():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=300)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--spinup", type=int, default=1500)
    p.add_argument("--workload", default="rough_bed", choices=["rough_bed", "plane_beach"])
    p.add_argument("--profile-steps", type=int, default=20)
    p.add_argument("--cpu-steps", type=int, default=4)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def cpu_baseline(w, state_arrays, t0, steps):
    """The CPU FP64 restatement (reference row order, glibc) on the host's cores, on `steps`
    steps from the same mid-run state; cell-updates/s."""
    from oracle import oracle as orc
    from wavesim_amd.grid import State

    cores = min(os.cpu_count() or 1, 16)
    out = {}
    for threads in (cores, 1):
        o = w.solver(lambda g, b: orc.OracleSolver(g, b, threads=threads))
        h, hu, hv, br = state_arrays
        s = State(h.copy(), hu.copy(), hv.copy(), br.copy(), t0)
        n = steps if threads > 1 else max(1, steps // 4)
        t = time.perf_counter()
        o.step_n(s, n)
        el = time.perf_counter() - t
        out[threads] = (w.grid.nx * w.grid.ny * n / el, n, el)
    return cores, out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "(DESIGN.md), not a measured path", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from wavesim_amd import configs
    from wavesim_amd.solver import Solver

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)
    torch.cuda.init()
    w = configs.c1_rough_bed() if args.workload == "rough_bed" else configs.c0_plane_beach()
    g = w.solver(Solver)
    s = w.still_water()
    n_cells = w.grid.nx * w.grid.ny
    print(f"[bench] {w.name}: spin-up {args.spinup} steps", file=sys.stderr, flush=True)
    t = time.perf_counter()
    done, _ = g.run_steps(s, args.spinup)
    print(f"[bench] spin-up {done} steps in {time.perf_counter() - t:.1f} s, t = {s.time:.1f} s",
          file=sys.stderr, flush=True)
    g.run_steps(s, args.warmup)
    st0 = g.dispersion_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done, last_dt = g.run_steps(s, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st1 = g.dispersion_stats()
    if done != args.steps:
        raise SystemExit(f"the run stopped after {done} steps (dt = {last_dt})")
    solves = st1.solves - st0.solves if st1 else 0
    iters = st1.iterations - st0.iterations if st1 else 0
    value = n_cells * args.steps / elapsed
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s: {value / 1e6:.1f} M cell-updates/s,"
          f" {iters / max(solves, 1):.1f} PCG iterations per solve", file=sys.stderr, flush=True)

    # Per-kernel device time, with HIP events on the solver's stream.
    g.profile(True)
    g.profile_reset()
    g.run_steps(s, args.profile_steps)
    prof = g.profile_read()
    g.profile(False)
    fam = {k: v for k, v in prof.items() if v[1] > 0}
    dominant = max(fam, key=lambda k: fam[k][0])
    tot_ms, launches = fam[dominant]
    avg_ms = tot_ms / launches
    bpc = BYTES_PER_CELL.get(dominant, 0)
    achieved = bpc * n_cells / (avg_ms * 1e-3) / 1e9
    kernels = {k: {"total_ms": round(v[0], 4), "launches": v[1],
                   "avg_us": round(1e3 * v[0] / v[1], 3)} for k, v in sorted(fam.items())}

    cpu = None
    if not args.no_cpu:
        arrays = [a.copy() for a in s.arrays()]
        cores, res = cpu_baseline(w, arrays, s.time, args.cpu_steps)
        v, n, el = res[cores]
        v1, n1, el1 = res[1]
        cpu = {"value": v, "unit": "cell-updates/s", "cores": cores, "kind": "port",
               "sample": f"{n} steps of the same workload from the same mid-run state "
                         f"({el:.1f} s), C++ FP64 restatement, reference row order and glibc, "
                         f"OpenMP rows",
               "value_1_thread": v1, "sample_1_thread": f"{n1} steps ({el1:.1f} s)"}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "nx": w.grid.nx, "ny": w.grid.ny, "dx_m": w.grid.dx,
                   "precision": "fp64 plain path", "pcg_tolerance": 1e-8,
                   "spinup_steps": args.spinup, "t_start_s": round(s.time, 3), **w.notes},
        "pcg_iters_per_solve": iters / max(solves, 1),
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_cell": bpc, "avg_launch_us": 1e3 * avg_ms},
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
