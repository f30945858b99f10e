"""Host-side checks of libvirial_lj.so that need no GPU: it loads, exports
every symbol include/virial_lj.h declares, validates arguments with the
reference's messages before touching a device, and its host lattices match
the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "virial_lj.h")


def header_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(vlj_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from virial_particles import _lib

    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 30
    for name in syms:
        assert hasattr(L, name), name
    assert set(syms) == set(_lib.SIGNATURES)


def test_create_validates_before_touching_a_device(lj):
    pos = np.zeros((10, 3))
    with pytest.raises(lj.LennardJonesError, match="at least two \\(cutoff \\+ skin\\) wide") as e:
        lj.LennardJones((pos, pos), lj.Settings([5.5, 20.0, 20.0], 2.5, 0.005))
    assert e.value.code == 1
    with pytest.raises(lj.LennardJonesError, match="Nose-Hoover needs .* positive period") as e:
        lj.LennardJones((pos, pos), lj.Settings([20.0] * 3, 2.5, 0.005, thermostat=lj.NoseHoover(1.0, 0.0)))
    assert e.value.code == 1
    with pytest.raises(lj.LennardJonesError, match="finite"):
        bad = pos.copy()
        bad[0, 0] = np.nan
        lj.LennardJones((bad, pos), lj.Settings([20.0] * 3, 2.5, 0.005))


def test_batch_replicas_share_precision(lj):
    pos = np.random.default_rng(1).uniform(0, 20, (10, 3))
    a = (pos, np.zeros_like(pos))
    with pytest.raises(lj.LennardJonesError, match="share"):
        lj.Batch([(a, lj.Settings([20.0] * 3, 2.5, 0.005)), (a, lj.Settings([20.0] * 3, 2.5, 0.005, precision=1))])


def test_no_device_is_a_loud_error(lj):
    from virial_particles import _lib

    n = C.c_int()
    dev = 0
    try:
        import torch  # noqa: F401 -- only to count devices

        dev = torch.cuda.device_count()
    except Exception:
        dev = 0
    if dev:
        pytest.skip("a GPU is present")
    pos = np.random.default_rng(1).uniform(0, 20, (10, 3))
    with pytest.raises(lj.LennardJonesError, match="no HIP device") as e:
        lj.LennardJones((pos, np.zeros_like(pos)), lj.Settings([20.0] * 3, 2.5, 0.005))
    assert e.value.code == _lib.E_NO_DEVICE
    del n


def test_host_lattices_match_the_oracle(oracle, lj):
    a = lj.lattice_array(1000, 11.0, 1.0, 3)
    p, v = oracle.lattice(1000, 11.0, 1.0, 3)
    assert np.array_equal(a["pos"], p) and np.array_equal(a["vel"], v)
    b = lj.lattice_box_array(1000, [9.0, 10.5, 14.25], 1.0, 3)
    p, v = oracle.lattice_box(1000, [9.0, 10.5, 14.25], 1.0, 3)
    assert np.array_equal(b["pos"], p) and np.array_equal(b["vel"], v)
    f, side = lj.fcc_array(5, 0.8442, 0.722, 7)
    p, v, s = oracle.fcc(5, 0.8442, 0.722, 7)
    assert np.array_equal(f["pos"], p) and np.array_equal(f["vel"], v) and side == s


def test_random_insertion_matches_the_oracle(oracle, lj):
    # the library's grid-accelerated insertion takes the same decisions as
    # the oracle's check against every earlier particle
    box = [12.0, 9.5, 14.0]
    a = lj.random_insertion_array(400, box, 0.9, 1.3, 11)
    p, v = oracle.random_insertion(400, box, 0.9, 1.3, 11)
    assert np.array_equal(a["pos"], p) and np.array_equal(a["vel"], v)
    pa = np.ascontiguousarray(a["pos"])
    d = pa[:, None, :] - pa[None, :, :]
    d -= np.array(box) * np.round(d / np.array(box))
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(400) * 1e9
    assert r.min() >= 0.9
    assert np.abs(a["vel"].sum(axis=0)).max() < 1e-9
    with pytest.raises(lj.LennardJonesError, match="no free place"):
        lj.random_insertion_array(200, [3.0, 3.0, 3.0], 0.9, 1.0, 1)


@pytest.mark.parametrize("count", [1, 500, 4 * 5 ** 3, 4 * 5 ** 3 + 1, 1201])
def test_fcc_for_any_count(lj, count):
    f, side = lj.fcc_count_array(count, 0.8442, 1.0, 5)
    assert side == float(np.float32((count / 0.8442) ** (1 / 3)))
    assert f.shape[0] == count
    assert (f["pos"] >= 0).all() and (f["pos"] < side).all()
    n = 1
    while 4 * n ** 3 < count:
        n += 1
    if count > 1:
        pf = np.ascontiguousarray(f["pos"])
        d = pf[:, None, :] - pf[None, :, :]
        d -= side * np.round(d / side)
        r = np.sqrt((d ** 2).sum(-1)) + np.eye(count) * 1e9
        assert r.min() == pytest.approx(side / n / np.sqrt(2), rel=1e-12)
        assert np.abs(f["vel"].sum(axis=0)).max() < 1e-9
    if count == 4 * 5 ** 3:  # a full lattice: the same sites as fcc(5), up to the side
        g, gside = lj.fcc_array(5, 0.8442, 1.0, 5)
        np.testing.assert_allclose(f["pos"] / side, g["pos"] / gside, rtol=1e-12, atol=1e-15)
        assert np.array_equal(f["vel"], g["vel"])


def test_lattice_api_mirror(lj):
    ps = lj.lattice(8, 4.0, 1.0, 1)
    assert len(ps) == 8 and isinstance(ps[0], lj.Particle)
    assert ps[0].pos == [1.0, 1.0, 1.0]


@pytest.mark.parametrize("n", [1, 7, 8, 9, 54, 127, 128, 129, 4096, 5407, 65536])
def test_table_dealing_covers_every_entry_once(n):
    from virial_particles import _lib

    f = _lib.lib().vlj_debug_swizzle
    # one round is 8 chunks; a one-entry table gets exactly one round
    rnd = f(1, 0)
    chunk = rnd // 8
    assert rnd == 8 * chunk and chunk == 4
    grid = f(n, 0)
    assert grid % rnd == 0 and n <= grid < n + rnd
    m = [f(b, 1) for b in range(grid)]
    assert sorted(m) == list(range(grid))
    # chunks of `chunk` consecutive entries, chunk c on XCD c % 8 (= b % 8)
    for b, idx in enumerate(m):
        assert (idx // chunk) % 8 == b % 8
    # a stretch of one round of consecutive entries touches all eight XCDs
    if n >= rnd:
        assert {b % 8 for b, idx in enumerate(m) if idx < rnd} == set(range(8))


def test_measurement_helpers(lj):
    m = lj.Measurement(0, 0.0, 3.0, -5.0, 2.0, [0, 0, 0], 10, 8.0, 1, 0, 0.25, 0)
    assert m.energy() == -2.0
    assert m.temperature(6.0) == 1.0
    assert m.pressure() == (6.0 + 2.0) / 24.0
    assert m.extended_energy() == -1.75


def test_load_rejects_files_that_are_not_states(lj, tmp_path):
    # vlj_load reads and checks the file before it touches a device
    with pytest.raises(lj.LennardJonesError, match="could not open"):
        lj.LennardJones.load(tmp_path / "missing.state")
    junk = tmp_path / "junk.state"
    junk.write_bytes(b"VLJSTATX" + bytes(200))
    with pytest.raises(lj.LennardJonesError, match="not a state file"):
        lj.LennardJones.load(junk)
    short = tmp_path / "short.state"
    short.write_bytes(b"VLJSTATE" + bytes(10))
    with pytest.raises(lj.LennardJonesError, match="not a state file|truncated"):
        lj.Batch.load(short)
