"""The reference's on-disk formats (crates/virial/src/data.rs,
manifest.rs): shortest exact numbers, .npy snapshots and run.json. The
reference's own tests (data.rs:184-213) plus NumPy as the reader."""
import io
import json
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "virial-engine_amd"))
from virial_particles import data  # noqa: E402


def test_numbers_round_trip_exactly():
    # data.rs:188-200
    for v in [0.1, -1.0 / 3.0, 6.02214076e23, 1e-300, -0.0, 12345.678901234567]:
        assert np.float64(float(data.exact(v))).tobytes() == np.float64(v).tobytes()
    rng = np.random.default_rng(1)
    for v in np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-30, 30, 2000),
                             rng.integers(-2**60, 2**60, 200).astype(np.float64)]):
        assert float(data.exact(v)) == v


def test_exact_writes_what_rust_debug_writes():
    # plain notation for 1e-4 <= |v| < 1e16, shortest exponential otherwise
    cases = [(0.1, "0.1"), (1.0, "1.0"), (-0.0, "-0.0"), (0.0, "0.0"), (1e15, "1000000000000000.0"),
             (1e16, "1e16"), (0.0001, "0.0001"), (0.00001, "1e-5"), (6.02214076e23, "6.02214076e23"),
             (1e-300, "1e-300"), (-2.5e-7, "-2.5e-7"), (123456.789, "123456.789"), (float("inf"), "inf"),
             (float("nan"), "NaN")]
    for v, text in cases:
        assert data.exact(v) == text


def test_display_is_plain():
    assert data.display(0.005) == "0.005"
    assert data.display(2.0) == "2"
    assert data.display(1e20) == "100000000000000000000"
    assert data.display(1e-7) == "0.0000001"
    assert data.display(1000) == "1000"
    assert data.display("2.5, shifted force") == "2.5, shifted force"


def test_npy_header_is_aligned_and_sized():
    # data.rs:202-212
    b = data.npy([3, 2], [1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    assert b[:8] == b"\x93NUMPY\x01\x00"
    header_len = int.from_bytes(b[8:10], "little")
    assert (10 + header_len) % 64 == 0
    assert len(b) == 10 + header_len + 6 * 8
    header = b[10:10 + header_len].decode()
    assert "'shape': (3, 2)" in header and header.endswith("\n")


@pytest.mark.parametrize("shape", [(5,), (3, 2), (2, 3, 4), (0,)])
def test_npy_is_numpys_own_format(shape):
    a = np.arange(int(np.prod(shape)), dtype=np.float64).reshape(shape) * 0.37 - 1.0
    b = data.npy(shape, a)
    ref = io.BytesIO()
    np.save(ref, a)
    assert b == ref.getvalue()
    np.testing.assert_array_equal(np.load(io.BytesIO(b)), a)
    with pytest.raises(ValueError):
        data.npy((7,), a)


def test_recorded_run_reads_back(tmp_path):
    log = data.DataLog(str(tmp_path))
    assert log.is_recording() and not data.DataLog().is_recording()
    vals = [(0.0, -5.503893613071572e6, 1.0 / 3.0), (0.005, 1e-300, 6.02214076e23)]
    for t, e, w in vals:
        log.row([("time", t), ("energy", e), ("virial", w)])
    with pytest.raises(ValueError):
        log.row([("time", 1.0), ("virial", 2.0)])
    log.meta("molecules", 1000)
    log.meta("dt", 0.005)
    log.meta("cutoff", "2.5, shifted force")
    log.meta("dt", 0.002)  # keeps its place, takes the new value
    log.snapshot("final_speeds", [4], [0.5, 1.5, 2.5, 3.5])
    log.save()
    data.write_manifest(log, title="gas", gpu="AMD Instinct MI355X", settings={"steps": "2000", "n": "1000"},
                        seconds=1.23456)
    cols, rows = data.read_csv(str(tmp_path / "measurements.csv"))
    assert cols == ["time", "energy", "virial"]
    assert [tuple(r) for r in rows] == vals
    np.testing.assert_array_equal(np.load(tmp_path / "final_speeds.npy"), [0.5, 1.5, 2.5, 3.5])
    text = (tmp_path / "run.json").read_text()
    run = json.loads(text)
    assert list(run) == ["scene", "virial_version", "git_commit", "git_uncommitted_changes", "gpu", "width",
                         "height", "fps", "frames", "render_seconds", "settings", "scene_meta"]
    assert run["render_seconds"] == 1.235 and run["git_commit"] is None
    assert list(run["settings"]) == ["n", "steps"]
    assert run["scene_meta"] == {"molecules": "1000", "dt": "0.002", "cutoff": "2.5, shifted force"}
    assert data.json_string('a"b\\c\n\x01') == '"a\\"b\\\\c\\n\\u0001"'
