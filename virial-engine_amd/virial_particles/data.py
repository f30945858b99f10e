"""The data a run writes, in the formats of the reference's app crate
(crates/virial/src/data.rs, manifest.rs): measurement rows to
`measurements.csv` with every number in its shortest exact form, array
snapshots to `<name>.npy` (NumPy format 1.0, little-endian f64) and the
run's facts to `run.json`. Byte for byte what the reference writes for the
same values, so its readers and this package's read each other's files.
"""
from __future__ import annotations

import json
import math
import os
import struct
from decimal import Decimal
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np


def exact(v: float) -> str:
    """The shortest decimal that reads back as exactly `v`, written the way
    the reference writes it (Rust's `{:?}` of an f64, data.rs:139-144):
    plain notation for 1e-4 <= |v| < 1e16 (with ".0" on whole numbers),
    otherwise `<digits>e<exponent>` with no '+' and no leading zeros."""
    v = float(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    r = repr(v)  # shortest round-trip digits, same thresholds for the notation
    if "e" in r:
        mant, exp = r.split("e")
        return f"{mant}e{int(exp)}"
    return r


def display(v) -> str:
    """A value as the reference's `meta` records it (Rust's `to_string`):
    floats in plain notation with their shortest exact digits and no
    trailing ".0", everything else by str()."""
    if isinstance(v, bool) or not isinstance(v, float):
        return str(v).lower() if isinstance(v, bool) else str(v)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    text = format(Decimal(repr(v)), "f")
    if "." in text:
        text = text.rstrip("0").rstrip(".")
    return "-0" if text == "-0" else text


def npy(shape: Sequence[int], data) -> bytes:
    """NumPy .npy (format 1.0) bytes of a little-endian f64 array, header
    padded so the data starts at a multiple of 64 (data.rs:147-167)."""
    a = np.ascontiguousarray(data, dtype="<f8").ravel()
    shape = tuple(int(d) for d in shape)
    if int(np.prod(shape)) != a.size:
        raise ValueError("shape doesn't match data")
    dims = [str(d) for d in shape]
    shape_text = f"({dims[0]},)" if len(dims) == 1 else f"({', '.join(dims)})"
    header = f"{{'descr': '<f8', 'fortran_order': False, 'shape': {shape_text}, }}"
    unpadded = 10 + len(header) + 1
    header += " " * ((64 - unpadded % 64) % 64) + "\n"
    return b"\x93NUMPY\x01\x00" + struct.pack("<H", len(header)) + header.encode() + a.tobytes()


def json_string(s: str) -> str:
    """A JSON string literal as the reference escapes it (data.rs:169-182)."""
    out = ['"']
    for c in s:
        if c == '"':
            out.append('\\"')
        elif c == "\\":
            out.append("\\\\")
        elif c == "\n":
            out.append("\\n")
        elif ord(c) < 0x20:
            out.append(f"\\u{ord(c):04x}")
        else:
            out.append(c)
    out.append('"')
    return "".join(out)


class DataLog:
    """Where a run's data goes (data.rs:14-128): rows, metadata and
    snapshots, kept in memory and written under `directory` when one is
    given (the reference's `record=<dir>`)."""

    def __init__(self, directory: Optional[str] = None):
        self.directory = directory
        self.columns: List[str] = []
        self.rows: List[List[float]] = []
        self.meta_entries: List[Tuple[str, str]] = []
        if directory is not None:
            os.makedirs(directory, exist_ok=True)

    def is_recording(self) -> bool:
        return self.directory is not None

    def row(self, values: Sequence[Tuple[str, float]]) -> None:
        """Adds a measurement row; the first row sets the columns and later
        rows must use the same names in the same order."""
        names = [n for n, _ in values]
        if not self.columns:
            self.columns = names
        elif names != self.columns:
            raise ValueError("measurement columns changed between rows")
        self.rows.append([float(v) for _, v in values])

    def meta(self, key: str, value) -> None:
        """Records a setting or fact for run.json; a key set again keeps its
        place and takes the new value."""
        value = value if isinstance(value, str) else display(value)
        for k, (name, _) in enumerate(self.meta_entries):
            if name == key:
                self.meta_entries[k] = (key, value)
                return
        self.meta_entries.append((key, value))

    def snapshot(self, name: str, shape: Sequence[int], data) -> None:
        """Writes `data` (row-major, `shape`) to <name>.npy when recording."""
        blob = npy(shape, data)
        if self.directory is not None:
            with open(os.path.join(self.directory, f"{name}.npy"), "wb") as f:
                f.write(blob)

    def csv(self) -> str:
        out = ",".join(self.columns) + "\n"
        for row in self.rows:
            out += ",".join(exact(v) for v in row) + "\n"
        return out

    def save(self) -> None:
        """Writes measurements.csv (when there are rows)."""
        if self.directory is not None and self.columns:
            with open(os.path.join(self.directory, "measurements.csv"), "w") as f:
                f.write(self.csv())


def manifest_json(data: DataLog, title: str, gpu: str, settings: Dict[str, str], width: int = 0, height: int = 0,
                  fps: int = 0, frames: int = 0, seconds: float = 0.0, commit: Optional[str] = None,
                  uncommitted_changes: Optional[bool] = None, version: str = "0.2.0") -> str:
    """run.json (manifest.rs:16-66): the run's facts in a fixed order, the
    settings sorted by key, the scene's own facts from `data.meta`."""
    lines = ["{"]

    def field(k: str, v: str) -> None:
        lines.append(f"  {json_string(k)}: {v},")

    field("scene", json_string(title))
    field("virial_version", json_string(version))
    field("git_commit", "null" if commit is None else json_string(commit))
    field("git_uncommitted_changes", "null" if uncommitted_changes is None else str(bool(uncommitted_changes)).lower())
    field("gpu", json_string(gpu))
    field("width", str(int(width)))
    field("height", str(int(height)))
    field("fps", str(int(fps)))
    field("frames", str(int(frames)))
    field("render_seconds", f"{seconds:.3f}")
    pairs = ", ".join(f"{json_string(k)}: {json_string(v)}" for k, v in sorted(settings.items()))
    field("settings", "{" + pairs + "}")
    meta = ", ".join(f"{json_string(k)}: {json_string(v)}" for k, v in data.meta_entries)
    lines.append(f'  "scene_meta": {{{meta}}}')
    lines.append("}")
    return "\n".join(lines) + "\n"


def write_manifest(data: DataLog, **kw) -> None:
    if data.directory is not None:
        with open(os.path.join(data.directory, "run.json"), "w") as f:
            f.write(manifest_json(data, **kw))


def read_csv(path: str) -> Tuple[List[str], np.ndarray]:
    """Columns and rows of a measurements.csv, every value exactly as written."""
    with open(path) as f:
        lines = [l.rstrip("\n") for l in f if l.strip()]
    cols = lines[0].split(",")
    rows = [[float(x) for x in l.split(",")] for l in lines[1:]]
    return cols, np.array(rows, dtype=np.float64).reshape(len(rows), len(cols))


def measurement_row(m, dof: Optional[float] = None) -> List[Tuple[str, float]]:
    """The columns a Lennard-Jones run records per measurement: time and the
    totals of lennard_jones.rs:84-131 with the derived temperature and
    pressure."""
    dof = 3.0 * m.count - 3.0 if dof is None else dof
    return [("time", m.time), ("kinetic", m.kinetic), ("potential", m.potential), ("energy", m.energy()),
            ("virial", m.virial), ("temperature", m.temperature(dof) if dof > 0 else 0.0),
            ("pressure", m.pressure()), ("thermostat_energy", m.thermostat_energy)]


def load_json(path: str) -> dict:
    with open(path) as f:
        return json.load(f)
