"""The run directory (crates/waveio/src/run.rs) through the library's native writer, on the CPU
(the writer's host entry points touch no GPU).

Pinned by the reference's own golden run (viewer/tests/fixtures/tiny, written by
crates/waveio/tests/golden.rs:17-79 and committed under tests/golden/run_tiny): the same run
must come out byte for byte. The rest mirrors crates/waveio/tests/formats.rs:149-260, and the
float layout of run.json is pinned by known answers of serde_json's printing (the ryu crate's
format64)."""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from wavesim_amd.run import Bed, Frame, Run, RunFormatError, RunWriter, read_header, remove_run

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "run_tiny")


def write_tiny(path) -> None:
    """golden.rs:17-79: a 6 x 4 bed rising 0.75 m per cell from -2 m, four frames of eta and
    breaking; every value a multiple of 1/16."""
    nx, ny = 6, 4
    n = np.arange(nx * ny)
    elevation = (np.float32(-2.0) + np.float32(0.75) * (n % nx).astype(np.float32))
    bed = Bed("tiny", nx, ny, 3.0, 2.0, elevation,
              Frame("EPSG:32604", 1000.0, 2000.0, 130.0, "fixture"),
              {"name": "Synthetic fixture", "attribution": "none"})
    waves = {"wave_height_m": 1.0, "wave_period_s": 14.0, "tide_m": 0.0, "maker_x_m": 6.0,
             "sponge_offshore_cells": 2, "sponge_side_cells": 1}
    w = RunWriter.create(path, bed, ["eta", "breaking"], waves)
    for k in range(4):
        wave = (np.float32(0.125) * np.float32(k) + np.float32(0.0625) *
                (n % nx).astype(np.float32) - np.float32(0.25))
        eta = np.where(elevation < 0, np.maximum(wave, elevation), elevation).astype(np.float32)
        breaking = np.where(elevation < 0, np.float32(0.25) * ((n + k) % 4), 0).astype(np.float32)
        w.write_frame(2.0 * k, [eta, breaking])
    w.finish()


def test_the_golden_run_comes_out_byte_for_byte(tmp_path):
    write_tiny(tmp_path)
    for name in ("run.json", "bed.f32", "frames.f32"):
        with open(os.path.join(GOLDEN, name), "rb") as f:
            want = f.read()
        with open(tmp_path / name, "rb") as f:
            got = f.read()
        assert got == want, name


def sample_bed(nx, ny):
    """formats.rs:40-46"""
    el = (np.float32(-5.0) + np.float32(0.01) * np.arange(nx * ny, dtype=np.float32))
    return Bed("sample", nx, ny, 3.0, 2.0, el)


def test_run_round_trips_frames_and_times(tmp_path):
    bed = sample_bed(6, 4)
    w = RunWriter.create(tmp_path, bed, ["eta"], {"period": 12.0})
    for k in range(3):
        w.write_frame(2.0 * k, [np.float32(k) + np.arange(24, dtype=np.float32) * 0.5])
    w.finish()
    run = Run.read(tmp_path)
    assert run.header["frame_count"] == 3
    assert run.header["times"] == [0.0, 2.0, 4.0]
    assert run.header["fields"] == ["eta"]
    assert run.header["waves"]["period"] == 12.0
    assert run.frame(2)[3] == 2.0 + 1.5
    assert os.path.getsize(tmp_path / "bed.f32") == 6 * 4 * 4


def test_a_run_written_over_an_old_one_has_no_header_until_it_finishes(tmp_path):
    bed = sample_bed(3, 2)
    w = RunWriter.create(tmp_path, bed, ["eta"], {})
    w.write_frame(0.0, [np.zeros(6)])
    w.write_stats(["eta_max"], [np.ones(6)], 0.0, 1.0)
    w.finish()
    (tmp_path / "notes.txt").write_text("mine")
    w = RunWriter.create(tmp_path, bed, ["eta"], {})
    w.write_frame(0.0, [np.full(6, 2.0)])
    del w  # dropped without finishing
    assert not (tmp_path / "run.json").exists(), "the old header survived"
    assert not (tmp_path / "stats.f32").exists(), "the old statistics survived"
    with pytest.raises(OSError):
        Run.read(tmp_path)
    assert (tmp_path / "notes.txt").read_text() == "mine"
    remove_run(tmp_path)
    assert sorted(os.listdir(tmp_path)) == ["notes.txt"]
    remove_run(tmp_path / "missing")  # a missing directory is not an error


def test_a_header_is_read_without_the_frames(tmp_path):
    bed = sample_bed(6, 4)
    w = RunWriter.create(tmp_path, bed, ["eta"], {"period": 12.0})
    w.write_frame(0.0, [np.zeros(24)])
    w.finish()
    os.remove(tmp_path / "frames.f32")
    h = read_header(tmp_path)
    assert (h["frame_count"], h["waves"]["period"]) == (1, 12.0)
    with pytest.raises(OSError):
        Run.read(tmp_path)


def test_a_frame_of_the_wrong_size_is_rejected(tmp_path):
    w = RunWriter.create(tmp_path, sample_bed(6, 4), ["eta"], None)
    with pytest.raises(RunFormatError, match="a field has 23 values, expected 24"):
        w.write_frame(0.0, [np.zeros(23)])


def test_a_run_can_carry_several_fields_per_frame(tmp_path):
    w = RunWriter.create(tmp_path, sample_bed(6, 4), ["eta", "breaking"], None)
    for k in range(3):
        eta = 10.0 * k + np.arange(24, dtype=np.float32)
        breaking = 0.5 + 100.0 * k + np.arange(24, dtype=np.float32)
        w.write_frame(float(k), [eta, breaking])
    w.finish()
    run = Run.read(tmp_path)
    assert run.header["fields"] == ["eta", "breaking"]
    assert run.frame(2)[5] == 25.0
    assert run.field("breaking", 2)[5] == 205.5
    assert run.field("breaking", 0)[0] == 0.5
    assert run.field("velocity", 0) is None
    assert os.path.getsize(tmp_path / "frames.f32") == 3 * 2 * 24 * 4


def test_the_wrong_number_of_fields_or_a_missing_eta_is_rejected(tmp_path):
    bed = sample_bed(6, 4)
    with pytest.raises(RunFormatError, match='must have an "eta" field'):
        RunWriter.create(tmp_path, bed, ["breaking"], None)
    w = RunWriter.create(tmp_path, bed, ["eta", "breaking"], None)
    with pytest.raises(RunFormatError, match="frame has 1 fields, expected 2"):
        w.write_frame(0.0, [np.zeros(24)])
    with pytest.raises(RunFormatError):
        w.write_frame(0.0, [np.zeros(24), np.zeros(23)])


def test_statistics_join_the_header(tmp_path):
    w = RunWriter.create(tmp_path, sample_bed(3, 2), ["eta"], None)
    w.write_frame(0.0, [np.zeros(6)])
    blocks = [np.arange(6, dtype=np.float32), -np.arange(6, dtype=np.float32)]
    w.write_stats(["eta_max", "eta_min"], blocks, 10.5, 66.0)
    with pytest.raises(RunFormatError, match="stats must be 2 blocks of 6 values"):
        w.write_stats(["eta_max", "eta_min"], [np.zeros(6), np.zeros(5)], 0.0, 1.0)
    path = w.finish()
    assert path.endswith("run.json")
    text = open(path).read()
    assert text.endswith('''  "stats": {
    "file": "stats.f32",
    "fields": [
      "eta_max",
      "eta_min"
    ],
    "from_s": 10.5,
    "to_s": 66.0
  }
}
''')
    run = Run.read(tmp_path)
    assert np.array_equal(run.stat("eta_min"), blocks[1]) and run.stat("wave_height") is None


# serde_json prints an f64 with the ryu crate's format64: the shortest digits that read back,
# laid out by where the decimal point falls.
FLOATS = [
    (1.0, "1.0"), (0.5, "0.5"), (3.0, "3.0"), (130.0, "130.0"), (1000.0, "1000.0"),
    (0.1, "0.1"), (123.456, "123.456"), (-2.75, "-2.75"), (0.0, "0.0"), (-0.0, "-0.0"),
    (1e-4, "0.0001"), (1e-5, "0.00001"), (2.5e-5, "0.000025"), (1e-6, "1e-6"),
    (1.23e-6, "1.23e-6"), (1e-7, "1e-7"), (1234567890123456.0, "1234567890123456.0"),
    (1e15, "1000000000000000.0"), (1e16, "1e16"), (1.2e16, "1.2e16"),
    (12345678901234567.0, "1.2345678901234568e16"), (5e-324, "5e-324"),
    (1.7976931348623157e308, "1.7976931348623157e308"), (0.30000000000000004,
                                                         "0.30000000000000004"),
    (float("nan"), "null"), (float("inf"), "null"),
]


def test_floats_print_as_serde_json_prints_them(tmp_path):
    w = RunWriter.create(tmp_path, sample_bed(3, 2), ["eta"], None)
    for k, (v, _) in enumerate(FLOATS):
        w.note(f"f{k:02d}", v)
    _, text = w.header()
    lines = [ln.rstrip(",") for ln in text.splitlines()]
    for k, (_, want) in enumerate(FLOATS):
        assert f'    "f{k:02d}": {want}' in lines, (k, want)


def test_values_keep_serde_jsons_layout(tmp_path):
    """Nested values against Python's json with the same layout (two-space indent, sorted
    keys, non-ASCII kept, the same escapes): for values without floats the two printers agree
    character for character."""
    rng = np.random.default_rng(5)
    alphabet = ['a', 'Z', ' ', '"', '\\', '\n', '\t', '\x01', '\x1f', '\x7f', 'é', '→', '𝄞', '/']

    def rand_value(depth):
        kind = rng.integers(0, 7 if depth < 3 else 4)
        if kind == 0:
            return int(rng.integers(-2**62, 2**62))
        if kind == 1:
            return "".join(rng.choice(alphabet) for _ in range(rng.integers(0, 6)))
        if kind == 2:
            return bool(rng.integers(0, 2))
        if kind == 3:
            return None
        if kind in (4, 5):
            keys = ["".join(rng.choice(alphabet) for _ in range(rng.integers(1, 4)))
                    for _ in range(rng.integers(0, 4))]
            return {k: rand_value(depth + 1) for k in keys}
        return [rand_value(depth + 1) for _ in range(rng.integers(0, 4))]

    for trial in range(40):
        v = {"k": rand_value(0), "z": [rand_value(1)], "a": {}}
        w = RunWriter.create(tmp_path, sample_bed(3, 2), ["eta"], v)
        _, text = w.header()
        want = json.dumps(v, indent=2, ensure_ascii=False, sort_keys=True)
        want = want.replace("\n", "\n  ")  # "waves" sits at the first level
        assert f'  "waves": {want},\n' in text, trial
        del w


def test_json_handed_over_is_checked(tmp_path):
    from wavesim_amd.run import rlib

    w = RunWriter.create(tmp_path, sample_bed(3, 2), ["eta"], None)
    for bad in (b"{", b"[1,]", b"01", b"1e999", b'"\\ud800"', b"nul", b"{} x"):
        assert rlib().wc_run_note(w._w, b"x", bad) != 0, bad
    assert rlib().wc_run_note(w._w, b"x", b' {"b": -0, "a": 18446744073709551616} ') == 0
    _, text = w.header()
    assert '"a": 1.8446744073709552e19' in text and '"b": -0.0' in text
    w.note("waves_recorded", 3)
    w.note("not_a_float", np.float64(2.0))
    _, text = w.header()
    assert '    "waves_recorded": 3,' in text and '    "not_a_float": 2.0' in text


def test_notes_make_waves_an_object(tmp_path):
    w = RunWriter.create(tmp_path, sample_bed(3, 2), ["eta"], [1, 2])
    w.note("truncated", True)
    w.note("stopped_at_s", 12.5)
    _, text = w.header()
    assert '  "waves": {\n    "stopped_at_s": 12.5,\n    "truncated": true\n  },' in text
    w = RunWriter.create(tmp_path, sample_bed(3, 2), ["eta"], None)
    _, text = w.header()
    assert '  "waves": null,\n  "source": null\n}' in text
