"""vlj_load checks everything in a state file that indexes device memory
before it touches a device: a file of the right sizes but with bad
contents gets VLJ_E_INVALID. The fixture is a state file the library wrote
(tests/golden/make_state_fixture.py: two liquid replicas after 30 Langevin
steps); each test corrupts one field and loads it. Without a GPU the intact
file passes every check and fails only at the device (VLJ_E_NO_DEVICE)."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "state_two_replicas.vljstate")
HEADER = struct.Struct("<8sII6Q3I5I")  # StateHeader
REPLICA_CELLS = 84  # offset of cells[3] in Replica (after box, half_box, inv_box, cell_size)


def sections(buf):
    (_, _, R, N, C_, list_total, prof_total, nb, _, _, _, hc_cap,
     sz_rep, sz_ctl, sz_set, sz_brick, n_dense) = HEADER.unpack_from(buf, 0)
    off = HEADER.size + R * sz_set
    rep_off = off
    nz = [struct.unpack_from("<3I", buf, rep_off + r * sz_rep + REPLICA_CELLS)[2] for r in range(R)]
    off += R * sz_rep
    depth = []  # per replica: prof_samples, then brick_x, brick_y, brick_z and sparse per z layer
    for r in range(R):
        depth.append(off + 8 + nz[r] * 8)
        off += 8 + nz[r] * 10
    s = {"R": R, "N": N, "C": C_, "nb": nb, "rep": rep_off, "sz_rep": sz_rep, "depth": depth,
         "sz_brick": sz_brick, "bricks": off}
    off += nb * sz_brick
    s["ctl"] = off
    off += R * sz_ctl
    off += 12 * N * 8
    for name, n, w in (("orig", N, 4), ("cell_of", N, 4), ("nbr_count", N, 4), ("cell_start", C_, 4),
                       ("cell_n", C_, 4), ("nbr", list_total, 2), ("cut", nb, 4), ("parts", nb * 64, 2),
                       ("cut_groups", (nb + 127) // 128, 4)):
        s[name] = off
        off += n * w
    off += prof_total * 8 + hc_cap * 3 * 8
    assert off == len(buf), "fixture layout"
    return s


def load(lj, path):
    from virial_particles import _lib

    h = C.c_void_p()
    rc = _lib.lib().vlj_load(os.fsencode(str(path)), C.byref(h))
    msg = _lib.lib().vlj_last_error(None).decode()
    if rc == 0:
        _lib.lib().vlj_destroy(h)
    return rc, msg


@pytest.fixture(scope="module")
def state():
    return bytearray(open(FIXTURE, "rb").read())


def corrupt(tmp_path, buf, edits):
    b = bytearray(buf)
    for at, fmt, value in edits:
        struct.pack_into(fmt, b, at, value)
    p = tmp_path / "bad.vljstate"
    p.write_bytes(bytes(b))
    return p


def test_the_fixture_passes_every_check(lj, state, tmp_path):
    from virial_particles import _lib

    p = tmp_path / "good.vljstate"
    p.write_bytes(bytes(state))
    rc, msg = load(lj, p)
    assert rc in (_lib.OK, _lib.E_NO_DEVICE), msg


def u32(buf, at):
    return struct.unpack_from("<I", buf, at)[0]


@pytest.mark.parametrize("case", [
    "brick_past_row", "bricks_overlap", "brick_depth", "layer_depth", "part_count", "bad_part", "cut_groups",
    "list_too_long",
    "cell_start", "cell_n", "cell_of", "orig_duplicate", "list_entry", "control_flag",
])
def test_corrupted_fields_are_rejected(lj, state, tmp_path, case):
    from virial_particles import _lib

    s = sections(state)
    nb, N = s["nb"], s["N"]
    b0, b1 = s["bricks"], s["bricks"] + s["sz_brick"]  # Brick: rep, cell0, gx, gy, gz
    edits = {
        "brick_past_row": [(b0 + 8, "<I", 8)],
        "bricks_overlap": [(b1 + 4, "<I", u32(state, b0 + 4)), (b1, "<I", u32(state, b0))],
        "brick_depth": [(b0 + 16, "<I", 3 - u32(state, b0 + 16))],
        "layer_depth": [(s["depth"][0], "<B", 3)],
        "part_count": [(s["cut"], "<I", 65)],
        # part code: x0 | y0 << 3 | (w - 1) << 6 | (h - 1) << 9 | z0 << 12; x0 7 and w 8 pass any brick's edge
        "bad_part": [(s["cut"], "<I", 1), (s["parts"], "<H", 7 | (0 << 3) | (7 << 6) | (0 << 9) | (0 << 12)),
                     (s["cut_groups"], "<I", 1)],
        "cut_groups": [(s["cut_groups"], "<I", 3)],
        "list_too_long": [(s["nbr_count"], "<I", 0xFFFF)],
        "cell_start": [(s["cell_start"], "<I", u32(state, s["cell_start"]) + 1)],
        "cell_n": [(s["cell_n"], "<I", u32(state, s["cell_n"]) + 1)],
        "cell_of": [(s["cell_of"], "<I", 0xFFFFFFF0)],
        "orig_duplicate": [(s["orig"] + 4, "<I", u32(state, s["orig"]))],
        "list_entry": [(s["nbr"], "<H", 0xFFF8)],
        "control_flag": [(s["ctl"] + 16, "<I", 7)],  # Control.rebuild
    }[case]
    assert nb > 1 and N > 1
    rc, msg = load(lj, corrupt(tmp_path, state, edits))
    assert rc == _lib.E_INVALID, f"{case}: {rc} {msg}"
