"""Throughput of wavecore's time step on one MI355X: cell-updates/s, PCG iterations per solve,
and the roofline of the dominant kernel, with the CPU FP64 path timed beside it.

    python bench.py [--steps K] [--warmup W] [--spinup S] [--workload large_reef] [--gpus 1]

Workload (default): the large single reef, 4096x2048 cells at 3 m (a 40 m shelf shoaling at
1:50 onto a 1 m reef flat, 8 Gaussian lobes, a 16-mode rough-bed overlay, a dry beach strip;
wave-maker H 3 m, T 16 s at 15 degrees, sponges W, S, N, Manning 0.025), fifth order, SGN by
PCG to 1e-8, breaking switch on. A step is one full Solver::step (both RK stages, every PCG
iteration) with dt = stable_dt, driven on the device by wc_run_steps.

The run: S - A spin-up steps in the FP32 form; from that state the FP64 path and the FP32 form
each take A more steps (A = --accuracy-steps), after which the FP32 form is compared with FP64
(surface, breaking mask, volume); then each of them goes on in its own solver (warm starts
full) with W untimed and K timed steps. The headline `value` is the FP64 path, the reference's
arithmetic, bit for bit the oracle's. Beside it in `paths`: the FP32 chosen form, the FP64
path with the extrapolated warm start, the FP32 form with the reference's warm start, and the
FP32 form with the learned preconditioner (these three from the FP64 state in solvers of their
own, primed until their warm start is full). Inputs are resident in HBM for every timed window.
`sweep` is the batched spot-sheet sweep (96 reef sheets of uneven size in one persistent launch
per phase), `cpu_baseline` the C++ FP64 restatement on the host's cores.
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

# Algorithmic HBM bytes per interior cell of one launch of each kernel family (each array read
# or written once, halo re-reads excluded), for the roofline.
BYTES_PER_CELL = {
    "fp64": {
        "cg_matvec": 8 * (2 + 2 + 2 + 1 + 1 + 2),  # read r, p_old, diag, c, h; write p
        # read p, c, h, r; write r; every other iteration also read w and p_{n-2}, write w
        "cg_update": 8 * (2 + 1 + 1 + 2 + 2) + 8 * (2 + 2 + 2) // 2,
        "cg_pair": 8 * (2 + 2 + 2 + 1 + 1 + 2) + 8 * (2 + 1 + 1 + 2 + 2) + 8 * (2 + 2 + 2) // 2,
        "cg_flush": 8 * (2 + 2 + 2),  # a solve ending after an odd iteration: read w, p; write w
        "cg_init": 8 * (2 + 1 + 1 + 2 + 2),  # read w, c, h, b; write r
        "sgn_setup": 8 * (7 + 5),  # read h, hu, hv, phi, lh, lhu, lhv; write c, diag, b
        "flux": 8 * (5 + 3),  # read h, hu, hv, b, phi; write lh, lhu, lhv
        "sgn_finish": 8 * (3 + 3 + 2 + 3 + 1 + 3),  # read s_in, rates, w, s, b; write s
        "switch": 8 * 3,  # read h, b; write phi
    },
    "fp32": {
        # one kernel per iteration: read r, Au, Ap, 1/diag, p (pairs), c, h; write p, Ap, r, Au
        # (pairs); the FP32 correction w (a pair) read and written every other iteration
        "cg_iter": 4 * (2 * 5 + 1 + 1 + 2 * 4) + 4 * 2,
        # init (read x0 and b FP64, w, c, h; write r, 1/diag, p, s) and start (read r,
        # 1/diag, c, h; write Au), the two launches of the family averaged
        "cg_init": (8 * 4 + 4 * (2 + 1 + 1) + 4 * 8 + 4 * (2 + 2 + 1 + 1) + 4 * 2) // 2,
        # read h, hu, hv, phi, lhu and lhv (FP64), lh (FP64); write c, b (FP64); warm start:
        # read x (FP64) and six past steps, write x0 (FP64), w = 0 and the step
        "sgn_setup": 4 * (3 + 1) + 8 * 3 + 4 + 8 * 2 + (8 * 2 + 4 * 2 * 6) + (8 * 2 + 4 * 2 + 4 * 2),
        # read the h, hu and hv pairs, phi, b (FP64); write lh, lhu, lhv (FP64)
        "flux": 4 * (2 + 2 + 2 + 1) + 8 + 8 * 3,
        # pair, s1, w, stage-1 momentum rates; l1h, lh, lhu, lhv, b (FP64); write pair; the
        # solution x (FP64) and its step read and written
        "sgn_finish": 4 * (6 + 4 + 2 + 2 + 6) + 8 * 5 + 2 * (8 * 2 + 4 * 2),
        "switch": 4 * 3 + 8,
    },
}
# Bytes per cell of one FP64 iteration of the reference's PCG as two fused passes (a matvec
# storing A p, an update reading it back): what the kernels here avoid moving.
TWO_PASS_FP64_PCG_BYTES = 192

# rocprofv3 --pmc summaries of the dominant kernels (traffic per launch)
TRAFFIC_PROFILES = [
    os.path.join(ROOT, "profiles", "cg_pair_large_reef_4096x2048_pmc.md"),
    os.path.join(ROOT, "profiles", "cg_iter_large_reef_4096x2048_pmc.md"),
    os.path.join(ROOT, "profiles", "flux_march_large_reef_4096x2048_pmc.md"),
]


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 --pmc summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a copy of known size), or None."""
    for path in TRAFFIC_PROFILES:
        try:
            with open(path) as f:
                for line in f:
                    if line.startswith(f"traffic_bytes_per_launch {kernel} "):
                        return float(line.split()[2])
        except OSError:
            pass
    return None


def parse():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--spinup", type=int, default=None,
                   help="steps before the window (default: 4000 for the large reef, by when its "
                        "waves break on the reef edge; 1500 otherwise), the last "
                        "--accuracy-steps of them in both arithmetics")
    p.add_argument("--accuracy-steps", type=int, default=1000,
                   help="steps the FP64 path and the FP32 form take from one state before they "
                        "are compared (and then timed)")
    p.add_argument("--workload", default="large_reef",
                   choices=["rough_bed", "plane_beach", "large_reef", "sweep"])
    p.add_argument("--graph", type=int, default=None, choices=[0, 1],
                   help="sweep: replay the fixed stretches of a step from HIP graphs (default 1)")
    p.add_argument("--sweep-size", default=None,
                   help="sweep: every instance at this size (e.g. 256x128) instead of the drawn ones")
    p.add_argument("--instances", type=int, default=96,
                   help="sweep: the number of reef-sheet instances (seeds 3000, 3001, ...)")
    p.add_argument("--sweep-spinup", type=int, default=300,
                   help="sweep leg of the default run: spin-up steps")
    p.add_argument("--sweep-steps", type=int, default=30,
                   help="sweep leg of the default run: timed steps (warm-up 5)")
    p.add_argument("--no-sweep", action="store_true", help="skip the sweep leg")
    p.add_argument("--profile-steps", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="CPU work per thread count for the baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--precision", default="both", choices=["fp64", "fp32", "both"])
    p.add_argument("--no-network", action="store_true",
                   help="skip the FP32 path with the learned preconditioner")
    p.add_argument("--no-breakdown", action="store_true",
                   help="skip the paths that split the FP32 form's gain")
    p.add_argument("--warm-start", type=int, default=None, choices=[0, 1, 2, 3, 4, 5],
                   help="FP32 form: order of the extrapolated warm start (0: the reference's "
                        "start from the previous solution)")
    p.add_argument("--warm-start-guard", type=int, default=None,
                   help="FP32 form: cells whose switch reads below this (per mille) start from "
                        "the previous solution")
    p.add_argument("--warm-start-window", type=int, default=None,
                   help="FP32 form: past solution steps the warm start weighs (>= its order)")
    p.add_argument("--pcg-lanes", type=int, default=None, choices=[256, 512],
                   help="blocks of the FP32 form's one-kernel PCG launches")
    p.add_argument("--pcg-march", type=int, default=None, choices=[0, 1],
                   help="one-kernel PCG layout: blocks march down column strips (1) or take "
                        "64x16 tiles (0, default)")
    p.add_argument("--pcg-walk", type=int, default=None,
                   help="one-kernel PCG on 64x16 tiles: tiles per block to aim for (ceil(tiles "
                        "/ n) blocks, at most the lanes; solver default 1, the sweep 8)")
    p.add_argument("--recon", type=int, default=None, choices=[32, 64],
                   help="FP32 form: the flux reconstruction's arithmetic (default 64)")
    p.add_argument("--pcg32", type=int, default=None,
                   help="PCG kernels of the FP32 form (3: two per iteration, 4: one)")
    p.add_argument("--pcg64", type=int, default=None,
                   help="PCG kernels of the FP64 path (3: 64x4 tiles; default 64x16 tiles)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a device copy of known size (to scale HBM counters) and exit")
    p.add_argument("--stream-ceilings", action="store_true",
                   help="only measure what plain HBM streams reach (copy, read, write, three "
                        "reads and a write; plain and nontemporal; 1 GiB arrays) and exit")
    return p.parse_args()


def calibrate_copy():
    """A plain device copy of 2 x 256 MiB, ten times, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": 4 * n, "copy_bytes_written_per_launch": 4 * n,
                      "launches": 10}))


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def configure(w, precision, args, *, network=None, warm_start=None, fp64_warm_start=None):
    """A solver of the workload in one of the measured configurations."""
    from wavesim_amd.solver import PRECISION_FP32, Solver

    g = w.solver(Solver)
    if precision == "fp64":
        if args.pcg64:
            g.set_option("pcg_kernels", args.pcg64)
        if fp64_warm_start:
            g.set_option("fp64_warm_start", fp64_warm_start)
            if args.warm_start_window is not None:
                g.set_option("warm_start_window", args.warm_start_window)
        return g
    g.with_precision(PRECISION_FP32)
    ws = args.warm_start if warm_start is None else warm_start
    if ws is not None:
        g.set_option("warm_start", ws)
    if args.warm_start_window is not None:
        g.set_option("warm_start_window", args.warm_start_window)
    if args.warm_start_guard is not None:
        g.set_option("warm_start_guard", args.warm_start_guard)
    if args.pcg32:
        g.set_option("pcg_kernels", args.pcg32)
    if args.pcg_lanes:
        g.set_option("pcg_lanes", args.pcg_lanes)
    if args.pcg_march is not None:
        g.set_option("pcg_march", args.pcg_march)
    if args.recon is not None:
        g.set_option("flux_reconstruction", args.recon)
    if network is not None:
        g.with_preconditioner(network)
    return g


def priming_steps(args, warm_start_order) -> int:
    """Steps a fresh solver needs before its warm start is what a running one has: the
    extrapolation's window of past steps (and one more), or one step for the reference's start
    from the previous solution."""
    if not warm_start_order:
        return 1
    window = args.warm_start_window if args.warm_start_window is not None else 6
    return window + 1


def timed_family(args, precision, network=False):
    """The PCG iteration family that profile mode 2 times in this configuration, or None."""
    if precision == "fp64":
        return "cg_pair" if args.pcg64 in (None, 4, 5) else None
    return "cg_iter" if (args.pcg32 in (None, 4) and not network) else None


def measure(w, g, s, precision, label, warmup, steps, profile_steps, prime=0, family=None):
    """Time `steps` steps of solver g from state s (after `warmup` untimed steps, of which
    `prime` beyond --warmup so that a fresh solver's warm start is full; the profile window
    shrinks by as many, so every path ends the same number of steps after its start). The PCG
    iterations are timed inside the timed window itself with one pair of HIP events on the
    solver's stream around each batch of back-to-back launches (FP32: the one-kernel iteration
    "cg_iter"; FP64: the matvec and update pair "cg_pair"); the kernel table comes from a pair
    of events around every launch over the profile steps after it."""
    import torch

    extra = max(0, prime - warmup)
    g.run_steps(s, warmup + extra)
    st0 = g.dispersion_stats()
    pcg_family = family
    one_kernel_or_pair = family is not None
    if one_kernel_or_pair:
        g.profile(2)
        g.profile_reset()
    torch.cuda.synchronize()
    t = time.perf_counter()
    done, last_dt = g.run_steps(s, steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    timed = g.profile_read() if one_kernel_or_pair else {}
    if one_kernel_or_pair:
        g.profile(0)
    st1 = g.dispersion_stats()
    if done != steps:
        raise SystemExit(f"{label}: the run stopped after {done} steps (dt = {last_dt})")
    solves = st1.solves - st0.solves if st1 else 0
    iters = st1.iterations - st0.iterations if st1 else 0
    n_cells = w.grid.nx * w.grid.ny
    g.profile(True)
    g.profile_reset()
    g.run_steps(s, max(profile_steps - extra, 1))
    prof = g.profile_read()
    g.profile(False)
    fam = {k: v for k, v in prof.items() if v[1] > 0}
    table = BYTES_PER_CELL[precision]
    data = {k: v for k, v in fam.items() if table.get(k, 0) > 0}
    if one_kernel_or_pair and timed.get(pcg_family, (0, 0))[1] > 0:
        dominant, window = pcg_family, "timed"
        tot_ms, launches = timed[pcg_family]
    else:
        dominant = max(data, key=lambda k: data[k][0])
        window = "profile"
        tot_ms, launches = data[dominant]
    avg_ms = tot_ms / launches
    # a launch after its solve has converged returns at once: count only the iterations done
    useful = min(iters, launches) if window == "timed" and iters > 0 else launches
    achieved = table[dominant] * n_cells * useful / (tot_ms * 1e-3) / 1e9
    kernels = {k: {"total_ms": round(v[0], 4), "launches": v[1],
                   "avg_us": round(1e3 * v[0] / v[1], 3),
                   "bytes_per_cell": table.get(k, 0)} for k, v in sorted(fam.items())}
    fallbacks = g.precond_fallbacks() if label.endswith("network") else 0
    log(f"{label}: {steps} steps in {elapsed:.3f} s = {n_cells * steps / elapsed / 1e6:.1f} M "
        f"cell-updates/s, {iters / max(solves, 1):.1f} PCG iterations per solve; {dominant} "
        f"{1e3 * tot_ms / max(useful, 1):.1f} us per iteration = {achieved:.0f} GB/s")
    return {
        "value": n_cells * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
        "pcg_iters_per_solve": iters / max(solves, 1), "dominant": dominant,
        "avg_ms": tot_ms / max(useful, 1), "achieved": achieved, "bytes_per_cell": table[dominant],
        "launches": launches, "iterations": iters, "window": window, "priming_steps": extra,
        "kernels": kernels, "solver": g, "state": s, "precond_fallbacks": fallbacks,
    }


def accuracy(w, g64, s64, g32, s32):
    """The FP32 form against FP64: surface error over wet cells, breaking masks (indicator >=
    0.8), volume."""
    grid = w.grid
    b = grid.inner(w.bed.b)
    h64, h32 = grid.inner(s64.arrays()[0]), grid.inner(s32.arrays()[0])
    wet = (h64 > 0.05) | (h32 > 0.05)
    d = np.where(wet, (h32 + b) - (h64 + b), 0.0)
    m64 = g64.breaking_indicator(s64) >= 0.8
    m32 = g32.breaking_indicator(s32) >= 0.8
    v64, v32 = g64.volume(s64), g32.volume(s32)
    return {"eta_max_err_m": float(np.max(np.abs(d))),
            "eta_rms_err_m": float(np.sqrt(np.mean(d[wet] ** 2))),
            "breaking_mask_agreement": float(np.mean(m64 == m32)),
            "breaking_cells_fp64": int(m64.sum()),
            "volume_rel_diff": abs(v32 - v64) / v64,
            "time_diff_s": abs(s32.time - s64.time)}


def cpu_rate(w, state_arrays, t0, seconds, threads):
    """The CPU FP64 restatement (reference row order, glibc) stepping from a state for about
    `seconds`; (cell-updates/s, steps, seconds)."""
    from oracle import oracle as orc
    from wavesim_amd.grid import State

    o = w.solver(lambda g, b: orc.OracleSolver(g, b, threads=threads))
    h, hu, hv, br = state_arrays
    s = State(h.copy(), hu.copy(), hv.copy(), br.copy(), t0)
    n, el = 0, 0.0
    t = time.perf_counter()
    while el < seconds and n < 10000:
        o.step_n(s, 1)
        n += 1
        el = time.perf_counter() - t
    return w.grid.nx * w.grid.ny * n / el, n, el


def cpu_baseline(w, state_arrays, t0, seconds):
    cores = min(os.cpu_count() or 1, 16)
    out = {}
    for threads in (cores, 1):
        out[threads] = cpu_rate(w, state_arrays, t0, seconds, threads)
        log(f"cpu baseline, {threads} threads: {out[threads][1]} steps in "
            f"{out[threads][2]:.1f} s")
    return cores, out


def sweep(args, spinup, warmup, steps, cpu_seconds):
    """The batched spot-sheet sweep: reef-sheet instances of uneven size (wavesim_amd.configs
    .c3_reef_instance, seeds 3000...), FP32 form, all stepped together by wc_batch_run_steps;
    then the same solvers one at a time for comparison. cell-updates/s counts every instance's
    interior cells. The PCG iteration launches (kb_cg1_iter) are timed over a few profiled
    steps after the timed window, each between its own events."""
    import torch

    from wavesim_amd import configs
    from wavesim_amd.solver import PRECISION_FP32, Batch, Solver

    size = tuple(int(v) for v in args.sweep_size.split("x")) if args.sweep_size else None
    ws = [configs.c3_reef_instance(seed, size=size)
          for seed in range(3000, 3000 + args.instances)]
    solvers = [w.solver(Solver).with_precision(PRECISION_FP32) for w in ws]
    walk = 8 if args.pcg_walk is None else args.pcg_walk
    for g in solvers:
        if args.pcg_lanes:
            g.set_option("pcg_lanes", args.pcg_lanes)
        g.set_option("pcg_tiles_per_block", walk)
    states = [w.still_water() for w in ws]
    cells = sum(w.grid.nx * w.grid.ny for w in ws)
    b = Batch(solvers)
    if args.graph is not None:
        b.set_option("graph", args.graph)
    log(f"sweep: {len(ws)} instances built, {cells / 1e6:.1f} M cells")
    t = time.perf_counter()
    for done in range(0, spinup, 100):  # in chunks, each reported
        b.run_steps(states, min(100, spinup - done))
        log(f"sweep spin-up: {min(done + 100, spinup)} steps, {time.perf_counter() - t:.1f} s")
    b.run_steps(states, warmup)
    st0 = [g.dispersion_stats() for g in solvers]
    torch.cuda.synchronize()
    t = time.perf_counter()
    done, _ = b.run_steps(states, steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    st1 = [g.dispersion_stats() for g in solvers]
    if min(done) != steps:
        raise SystemExit(f"sweep: an instance stopped after {min(done)} steps")
    iters = sum(b1.iterations - b0.iterations for b0, b1 in zip(st0, st1))
    solves = sum(b1.solves - b0.solves for b0, b1 in zip(st0, st1))
    value = cells * steps / elapsed
    log(f"sweep batched: {steps} steps in {elapsed:.3f} s = {value / 1e6:.1f} M cell-updates/s, "
        f"{iters / max(solves, 1):.1f} PCG iterations per solve")
    # the PCG iteration launches, timed one by one over a few more steps
    prof_steps = 3
    p0 = [g.dispersion_stats() for g in solvers]
    b.set_option("profile", 1)
    b.run_steps(states, prof_steps)
    ms, launches = b.profile_read()
    b.set_option("profile", 0)
    p1 = [g.dispersion_stats() for g in solvers]
    cell_iters = sum(w.grid.nx * w.grid.ny * (b1.iterations - b0.iterations)
                     for w, b0, b1 in zip(ws, p0, p1))
    per_cell = BYTES_PER_CELL["fp32"]["cg_iter"]
    achieved = per_cell * cell_iters / (ms * 1e-3) / 1e9
    alone_steps = max(steps // 5, 2)
    for g in solvers:  # alone, a small grid is fastest with a PCG block per tile
        g.set_option("pcg_tiles_per_block", 1)
    for g, s in zip(solvers, states):  # untimed: each solver's first launches alone (the flux
        g.run_steps(s, 1)              # band height is tuned at the first)
    st0 = [g.dispersion_stats() for g in solvers]
    torch.cuda.synchronize()
    t = time.perf_counter()
    for g, s in zip(solvers, states):
        g.run_steps(s, alone_steps)
    torch.cuda.synchronize()
    el1 = time.perf_counter() - t
    st1 = [g.dispersion_stats() for g in solvers]
    iters1 = sum(b1.iterations - b0.iterations for b0, b1 in zip(st0, st1))
    solves1 = sum(b1.solves - b0.solves for b0, b1 in zip(st0, st1))
    v1 = cells * alone_steps / el1
    log(f"sweep one at a time: {alone_steps} steps in {el1:.3f} s = {v1 / 1e6:.1f} M "
        f"cell-updates/s, {iters1 / max(solves1, 1):.1f} PCG iterations per solve")
    sizes = {}
    for w in ws:
        sizes[f"{w.grid.nx}x{w.grid.ny}"] = sizes.get(f"{w.grid.nx}x{w.grid.ny}", 0) + 1
    cpu = None
    if cpu_seconds > 0:
        # a sample of the same workload: the sweep's first 512x256 instance from its state
        q = next((i for i, w in enumerate(ws) if (w.grid.nx, w.grid.ny) == (512, 256)), 0)
        cores = min(os.cpu_count() or 1, 16)
        arrays = [a.copy() for a in states[q].arrays()]
        v, n, el = cpu_rate(ws[q], arrays, states[q].time, cpu_seconds, cores)
        cpu = {"value": v, "unit": "cell-updates/s", "cores": cores, "kind": "port",
               "sample": f"{n} steps of sweep instance {q} ({ws[q].grid.nx}x{ws[q].grid.ny}, seed "
                         f"{3000 + q}) from its spun-up state ({el:.1f} s), C++ FP64 "
                         f"restatement, reference row order and glibc, OpenMP rows"}
        log(f"sweep cpu baseline: {n} steps in {el:.1f} s")
    return {
        "value": value, "unit": "cell-updates/s", "ms_per_step": 1e3 * elapsed / steps,
        "steps": steps, "warmup": warmup, "dtype": "f32",
        "config": {"workload": f"reef_sheet_sweep_{len(ws)}" + (f"_{args.sweep_size}"
                                                                if args.sweep_size else ""),
                   "instances": len(ws), "cells": cells, "sizes": sizes, "spinup_steps": spinup,
                   "path": "FP32 form, one launch per phase over all instances",
                   "pcg_tiles_per_block": walk, "launch_widths": list(b.launch_widths()),
                   "flux_band_rows": b.flux_rows()[0], "setup_tile_rows": b.setup_rows()[0]},
        "pcg_iters_per_solve": iters / max(solves, 1),
        "roofline": {"bound": "hbm", "kernel": "kb_cg1_iter", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_cell": per_cell,
                     "cell_iterations": cell_iters, "launches": launches,
                     "avg_launch_us": 1e3 * ms / max(launches, 1),
                     "window": f"{prof_steps} profiled steps after the timed window"},
        "one_at_a_time": {"value": v1, "steps": alone_steps, "batched_over_alone": value / v1,
                          "pcg_iters_per_solve": iters1 / max(solves1, 1)},
        "cpu_baseline": cpu,
    }


def main():
    args = parse()
    if args.spinup is None:
        args.spinup = 4000 if args.workload == "large_reef" else 1500
    if args.gpus != 1:
        print(json.dumps({"error": "wavesim_amd steps on one GPU; sharding over several "
                                   "is designed, not built or measured", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from wavesim_amd import configs

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        sys.exit(1)
    torch.cuda.init()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.stream_ceilings:
        from wavesim_amd.solver import stream_ceilings

        for grid in (2048, 4096, 8192, 16384):
            print(json.dumps({"stream_tb_per_s": stream_ceilings(1 << 30, grid, 10),
                              "bytes_per_array": 1 << 30, "grid": grid, "threads": 256,
                              "reps": 10}))
        return
    if args.workload == "sweep":
        out = sweep(args, args.spinup, args.warmup, args.steps,
                    0.0 if args.no_cpu else args.cpu_seconds)
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "data": "synthetic", **out},
                         default=float))
        return
    w = {"rough_bed": configs.c1_rough_bed, "plane_beach": configs.c0_plane_beach,
         "large_reef": configs.c4_large_reef}[args.workload]()
    from wavesim_amd.grid import State

    acc_steps = min(args.accuracy_steps, args.spinup)
    # 1. the FP32 form to the start of the accuracy window
    g32 = configure(w, "fp32", args)
    s32 = w.still_water()
    log(f"{w.name}: spin-up {args.spinup - acc_steps} steps (FP32 form)")
    t = time.perf_counter()
    g32.run_steps(s32, args.spinup - acc_steps)
    log(f"spin-up in {time.perf_counter() - t:.1f} s, t = {s32.time:.1f} s")
    # 2. both arithmetics from that one state through the accuracy window
    h, hu, hv, br = (a.copy() for a in s32.arrays())
    s64 = State(h, hu, hv, br, s32.time)
    g64 = configure(w, "fp64", args) if args.precision != "fp32" else None
    t = time.perf_counter()
    if g64 is not None:
        g64.run_steps(s64, acc_steps)
        log(f"FP64 path through the {acc_steps}-step window in {time.perf_counter() - t:.1f} s")
    g32.run_steps(s32, acc_steps)
    window_acc = accuracy(w, g64, s64, g32, s32) if g64 is not None else None
    if window_acc:
        log(f"FP32 form vs FP64 after {acc_steps} steps: {window_acc}")
    # every further path starts from the FP64 state (the FP32 form from its own)
    base = s64 if g64 is not None else s32
    arrays = [a.copy() for a in base.arrays()]
    t0 = base.time

    def fresh_state():
        return State(*(a.copy() for a in arrays), t0)

    res = {}
    if g64 is not None:
        res["fp64"] = measure(w, g64, s64, "fp64", "fp64", args.warmup, args.steps,
                              args.profile_steps, family=timed_family(args, "fp64"))
    if args.precision != "fp64":
        res["fp32"] = measure(w, g32, s32, "fp32", "fp32", args.warmup, args.steps,
                              args.profile_steps, family=timed_family(args, "fp32"))
    if not args.no_breakdown and g64 is not None:
        # the FP64 path with the extrapolated warm start: what the start gains without FP32
        g = configure(w, "fp64", args, fp64_warm_start=3)
        res["fp64+extrapolated_start"] = measure(
            w, g, fresh_state(), "fp64", "fp64+extrapolated_start", args.warmup, args.steps,
            args.profile_steps, prime=priming_steps(args, 3), family=timed_family(args, "fp64"))
        if args.precision != "fp64":
            # the FP32 form with the reference's warm start: what storage alone gains
            g = configure(w, "fp32", args, warm_start=0)
            res["fp32+previous_start"] = measure(
                w, g, fresh_state(), "fp32", "fp32+previous_start", args.warmup, args.steps,
                args.profile_steps, prime=priming_steps(args, 0),
                family=timed_family(args, "fp32"))
    # the network trained on the residuals the FP32 form's extrapolated starts leave
    net_file = os.path.join(ROOT, "wavesim_amd", "data", "precond_extrapolated.npy")
    if args.precision != "fp64" and not args.no_network and os.path.exists(net_file):
        g = configure(w, "fp32", args, network=np.load(net_file))
        order = 3 if args.warm_start is None else args.warm_start
        res["fp32+network"] = measure(w, g, fresh_state(), "fp32", "fp32+network", args.warmup,
                                      args.steps, args.profile_steps,
                                      prime=priming_steps(args, order),
                                      family=timed_family(args, "fp32", network=True))
    accs = {}
    if "fp64" in res:
        r64 = res["fp64"]
        for p, r in res.items():
            if p != "fp64":
                accs[p] = accuracy(w, r64["solver"], r64["state"], r["solver"], r["state"])
    head = "fp64" if "fp64" in res else "fp32"
    r = res[head]
    cpu = None
    if not args.no_cpu:
        cores, cres = cpu_baseline(w, arrays, t0, args.cpu_seconds)
        v, n, el = cres[cores]
        v1, n1, el1 = cres[1]
        cpu = {"value": v, "unit": "cell-updates/s", "cores": cores, "kind": "port",
               "sample": f"{n} steps of the same workload from the FP64 path's window state "
                         f"({el:.1f} s), C++ FP64 restatement, reference row order and glibc, "
                         f"OpenMP rows",
               "value_1_thread": v1, "sample_1_thread": f"{n1} steps ({el1:.1f} s)"}
    sw = None
    if args.workload == "large_reef" and not args.no_sweep:
        sw = sweep(args, args.sweep_spinup, 5, args.sweep_steps,
                   0.0 if args.no_cpu else min(args.cpu_seconds, 10.0))

    ratio = {p: v["value"] / res["fp64"]["value"] for p, v in res.items()} \
        if "fp64" in res else {}
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "nx": w.grid.nx, "ny": w.grid.ny, "dx_m": w.grid.dx,
                   "path": ("FP64 plain path: the reference's arithmetic and PCG (Jacobi, warm "
                            "start from the previous solution), bit for bit the oracle's"
                            if head == "fp64" else "FP32 chosen form"),
                   "pcg_tolerance": 1e-8,
                   "spinup_steps": args.spinup, "accuracy_window_steps": acc_steps,
                   "t_start_s": round(t0, 3), **w.notes},
        "pcg_iters_per_solve": r["pcg_iters_per_solve"],
        "roofline": {"bound": "hbm", "kernel": r["dominant"], "achieved": r["achieved"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["achieved"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(r["dominant"]),
                     "algorithmic_bytes_per_launch": r["bytes_per_cell"] * w.grid.nx * w.grid.ny,
                     "bytes_per_cell": r["bytes_per_cell"],
                     **({"two_pass_bytes_per_cell": TWO_PASS_FP64_PCG_BYTES,
                         "kernels": "k_cg2_matvec + k_cg2_update (one PCG iteration)"}
                        if r["dominant"] == "cg_pair" else {}),
                     "avg_launch_us": 1e3 * r["avg_ms"], "launches": r["launches"],
                     "pcg_iterations": r["iterations"], "window": r["window"]},
        "paths": {p: {"value": v["value"], "ms_per_step": v["ms_per_step"],
                      "pcg_iters_per_solve": v["pcg_iters_per_solve"],
                      "vs_fp64_speed": ratio.get(p), "priming_steps": v["priming_steps"],
                      "precond_fallbacks": v["precond_fallbacks"],
                      "pcg_us_per_iteration": 1e3 * v["avg_ms"] if v["window"] == "timed"
                      else None,
                      "kernels": v["kernels"], "vs_fp64": accs.get(p)}
                  for p, v in res.items()},
        "fp32_vs_fp64": {"window_steps": acc_steps + args.warmup + args.steps
                         + args.profile_steps, **accs["fp32"]} if "fp32" in accs else window_acc,
        "speedup_target": {"target": "FP32 + learned preconditioner >= 4x the FP64 path",
                       "fp32+network_over_fp64": ratio.get("fp32+network"),
                       "fp32_over_fp64": ratio.get("fp32")},
        "cpu_baseline": cpu,
        "sweep": sw,
    }
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
