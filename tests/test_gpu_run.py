"""Frames and statistics written from the GPU into a run directory (wc_run_frame_from,
wc_run_stats_from): Stage::frame_if_due's frame (crates/wavesim/src/lib.rs:1348-1380) and
Stage::finish's statistics (lib.rs:1436-1509) against the same quantities from the oracle, in
lockstep on the FP64 path, where the GPU reproduces the oracle bit for bit. Every frame goes
through the writer's pinned slots and its own thread; the file must hold the frames in the
order they were taken."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from wavesim_amd import configs
from wavesim_amd.grid import State
from wavesim_amd.run import Bed, Run, RunWriter
from wavesim_amd.solver import STAGE_OK, Solver, Stats, stage_step

pytestmark = pytest.mark.gpu


def lockstep(nx, ny, seed, period, stats_from, until, frame_every, tmp_path, tide=0.0):
    """Runs the GPU and the oracle side by side; the GPU writes its frames into a run, the
    oracle's frames are computed on the host from its state."""
    w = configs.c1_rough_bed(nx=nx, ny=ny, seed=seed)
    g = w.solver(Solver)
    o = w.solver(lambda gr, b: orc.OracleSolver(gr, b, dot_order=orc.DOT_CANONICAL16,
                                                math=orc.MATH_PORTABLE))
    s = w.still_water()
    s.time = 20.0
    so = s.clone()
    elevation = w.grid.inner(w.bed.b).astype(np.float32).ravel() + np.float32(0.25)
    bed = Bed("rough", nx, ny, w.grid.dx, w.grid.dy, elevation)
    writer = RunWriter.create(tmp_path, bed, ["eta", "breaking"], {"wave_period_s": period})
    sg = Stats(g, stats_from, period, elevation)
    st = orc.OracleStats(o, stats_from, period, elevation)
    want, rises = [], []
    next_frame = s.time
    while s.time < until - 1e-9:
        if s.time >= next_frame - 1e-9:
            ok, rise = writer.frame_from(g, s, tide)
            assert ok
            inner = w.grid.inner(so.arrays()[0]).ravel()
            eta = (inner + elevation.astype(np.float64)).astype(np.float32)
            brk = o.breaking_indicator(so).astype(np.float32)
            want.append((so.time, eta, brk))
            deep = inner > 1e-3
            rises.append(max(0.0, float(np.max(eta[deep].astype(np.float64) - tide))))
            assert rise == rises[-1]
            next_frame += frame_every
        dt, status, _ = stage_step(g, s, sg, None, until)
        assert status == STAGE_OK
        do = min(o.stable_dt(so), until - so.time)
        assert dt == do
        o.step(so, do)
        st.record(so, do)
    writer.stats_from(sg)
    writer.finish()
    return w, Run.read(tmp_path), want, sg, st, s


def test_frames_and_statistics_from_the_gpu_match_the_oracle(tmp_path):
    period = 14.0
    w, run, want, sg, st, s = lockstep(96, 48, 11, period, 22.0, 20.0 + 3.2 * period, 2.0,
                                       tmp_path)
    assert run.header["frame_count"] == len(want) >= 20
    assert run.header["times"] == [t for t, _, _ in want]
    for k, (_, eta, brk) in enumerate(want):
        assert np.array_equal(run.field("eta", k), eta), k
        assert np.array_equal(run.field("breaking", k), brk), k
    # Stage::finish: close the wave that is due, then the blocks as f32
    st.finish()
    ro = st.read()
    n = w.grid.nx * w.grid.ny
    hdr = run.header["stats"]
    assert hdr["from_s"] == 22.0 and hdr["to_s"] == s.time
    assert ro["waves"] >= 3
    assert hdr["fields"] == ["break_fraction", "eta_max", "eta_min", "wave_height",
                             "harmonic_re", "harmonic_im"]
    frac = (ro["breaking_seconds"] / max(ro["seconds"], np.finfo(np.float64).tiny))
    assert np.array_equal(run.stat("break_fraction"), frac.astype(np.float32))
    assert np.array_equal(run.stat("eta_max"), ro["eta_max"].astype(np.float32))
    assert np.array_equal(run.stat("eta_min"), ro["eta_min"].astype(np.float32))
    assert np.array_equal(run.stat("wave_height"), ro["wave_mean"].astype(np.float32))
    a, b = st.harmonic()
    assert np.array_equal(run.stat("harmonic_re"), a.astype(np.float32))
    assert np.array_equal(run.stat("harmonic_im"), b.astype(np.float32))
    assert run.stats.size == 6 * n
    assert run.header["waves"]["waves_recorded"] == ro["waves"]
    assert "unsettled_per_period" in run.header["waves"]


def test_an_unstable_surface_is_not_written(tmp_path):
    w = configs.c1_rough_bed(nx=64, ny=32, seed=3)
    g = w.solver(Solver)
    s = w.still_water()
    elevation = w.grid.inner(w.bed.b).astype(np.float32).ravel()
    bed = Bed("rough", 64, 32, w.grid.dx, w.grid.dy, elevation)
    writer = RunWriter.create(tmp_path, bed, ["breaking", "eta"], None)
    g.run_steps(s, 3)
    ok, rise = writer.frame_from(g, s)
    assert ok and 0.0 <= rise < 1.0
    h = s.arrays()[0].copy()
    k = w.grid.idx(10, 10)
    tall = h.copy()
    tall[k] += 150.0  # a surface 150 m up: not a frame
    g2 = w.solver(Solver)
    s2 = State(tall, *[a.copy() for a in s.arrays()[1:]], s.time)
    ok, rise = writer.frame_from(g2, s2)
    assert not ok and rise > 100.0
    nan = h.copy()
    nan[k] = np.nan
    s3 = State(nan, *[a.copy() for a in s.arrays()[1:]], s.time)
    ok, _ = writer.frame_from(g2, s3)
    assert not ok
    writer.frame_from(g, s)
    writer.finish()
    run = Run.read(tmp_path)
    assert run.header["frame_count"] == 2  # the two refused frames left nothing behind
    assert np.array_equal(run.field("eta", 0), run.field("eta", 1))


def test_device_and_host_frames_keep_their_order_at_full_size(tmp_path):
    """1024 x 512 cells (the rough-bed workload's grid): frames from the device and from the
    host interleaved, more than the writer has slots, each checked against the state it was
    taken from."""
    w = configs.c1_rough_bed()
    g = w.solver(Solver)
    s = w.still_water()
    elevation = w.grid.inner(w.bed.b).astype(np.float32).ravel()
    bed = Bed("rough", w.grid.nx, w.grid.ny, w.grid.dx, w.grid.dy, elevation)
    writer = RunWriter.create(tmp_path, bed, ["eta", "breaking"], None)
    want = []
    for k in range(8):
        g.run_steps(s, 5)
        if k % 3 == 2:
            eta = np.full(elevation.size, np.float32(k))
            writer.write_frame(s.time, [eta, -eta])
            want.append((eta, -eta))
        else:
            writer.frame_from(g, s)
            h = w.grid.inner(s.arrays()[0]).ravel()  # downloads the state
            want.append(((h + elevation.astype(np.float64)).astype(np.float32),
                         g.breaking_indicator(s).astype(np.float32)))
    writer.finish()
    run = Run.read(tmp_path)
    assert run.header["frame_count"] == 8
    for k, (eta, brk) in enumerate(want):
        assert np.array_equal(run.field("eta", k), eta), k
        assert np.array_equal(run.field("breaking", k), brk), k
