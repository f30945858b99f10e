"""The oracle's profiles along z (lennard_jones.wgsl:628-730), pinned by
closed forms for one pair (including a pair across the periodic boundary)
and by sums that the profiles must add up to: particle count, kinetic
energy, potential energy and virial of the measurement, and the brute-force
all-pairs virial tensor."""
import numpy as np
import pytest

from helpers import sc_liquid

TWO32 = 2.0 ** 32
RC = 2.5


def lj_terms(r2, rc=RC):
    """r.F and U of the truncated and shifted pair at r2, plain formulas."""
    s6 = 1.0 / r2 ** 3
    sc6 = 1.0 / rc ** 6
    return 48.0 * s6 * s6 - 24.0 * s6, 4.0 * (s6 * s6 - s6) - 4.0 * (sc6 * sc6 - sc6)


def pair_profile(oracle, za, zb, dy, bins=12, side=12.0):
    pos = np.array([[5.0, 5.0, za], [5.0, 5.0 + dy, zb]])
    sim = oracle.Sim(pos, np.zeros((2, 3)), [side] * 3)
    sim.set_profile_bins(bins)
    sim.sample_profile()
    raw, samples = sim.read_profile()
    assert samples == 1
    return raw.reshape(-1)[: bins * 8].reshape(bins, 8) / TWO32, raw[bins * 8:] / TWO32


def test_one_pair_inside_the_box(oracle):
    slabs, total = pair_profile(oracle, 1.2, 2.4, 0.3)
    d = np.array([0.0, 0.3, 1.2])
    r2 = float(d @ d)
    rf, u = lj_terms(r2)
    w = d * d * rf / r2            # the whole pair's virial tensor diagonal
    # Counts and half the energy in each particle's slab
    assert slabs[1, 0] == 1.0 and slabs[2, 0] == 1.0 and slabs[:, 0].sum() == 2.0
    np.testing.assert_allclose(slabs[[1, 2], 7], [0.5 * u, 0.5 * u], rtol=1e-9)
    # The line from z = 1.2 to 2.4 lies 0.8 in slab 1 and 0.4 in slab 2
    np.testing.assert_allclose(slabs[1, 4:7], w * 0.8 / 1.2, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(slabs[2, 4:7], w * 0.4 / 1.2, rtol=1e-9, atol=1e-9)
    assert np.all(slabs[[0] + list(range(3, 12)), 4:] == 0.0)
    np.testing.assert_allclose(total, w, rtol=1e-9, atol=1e-9)


def test_one_pair_across_the_boundary(oracle):
    # z = 0.3 and 11.5 in a box 12 high: the line runs from -0.5 to 0.3,
    # 0.5 of it in the top slab (11) and 0.3 in the bottom one (0)
    slabs, total = pair_profile(oracle, 0.3, 11.5, 0.4)
    d = np.array([0.0, 0.4, -0.8])
    r2 = float(d @ d)
    rf, _ = lj_terms(r2)
    w = d * d * rf / r2
    np.testing.assert_allclose(slabs[11, 4:7], w * 0.5 / 0.8, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(slabs[0, 4:7], w * 0.3 / 0.8, rtol=1e-9, atol=1e-9)
    assert np.all(slabs[1:11, 4:7] == 0.0)
    assert slabs[0, 0] == 1.0 and slabs[11, 0] == 1.0


def test_one_pair_within_one_slab(oracle):
    # Both in slab 3: all of the virial stays there, the two halves adding up
    slabs, total = pair_profile(oracle, 3.2, 3.7, 1.0)
    d = np.array([0.0, 1.0, 0.5])
    rf, _ = lj_terms(float(d @ d))
    np.testing.assert_allclose(slabs[3, 4:7], d * d * rf / float(d @ d), rtol=1e-9, atol=1e-9)
    assert np.all(np.delete(slabs, 3, axis=0)[:, 4:7] == 0.0)


def test_pair_in_one_plane_goes_to_its_slab(oracle):
    # Equal z: the line has no length along z and the whole share goes home
    slabs, _ = pair_profile(oracle, 5.5, 5.5, 1.1)
    rf, _ = lj_terms(1.21)
    assert slabs[5, 5] == pytest.approx(rf, rel=1e-9)
    assert slabs[5, 6] == 0.0 and slabs[5, 4] == 0.0


@pytest.mark.parametrize("bins", [1, 7, 40])
def test_profiles_add_up_to_the_measurement(oracle, bins):
    pos, vel, box = sc_liquid(oracle, n_side=10)
    sim = oracle.Sim(oracle.jitter(pos, 5, 0.3), vel, [box] * 3)
    sim.set_profile_bins(bins)
    sim.run(30)
    sim.sample_profile()
    raw, samples = sim.read_profile()
    m = sim.measure()
    slabs = raw[: bins * 8].reshape(bins, 8)
    assert samples == 1
    assert slabs[:, 0].sum() == sim.n * 2 ** 32
    _, v = sim.read_particles()
    np.testing.assert_allclose(slabs[:, 1:4].sum(axis=0) / TWO32, (v * v).sum(axis=0), rtol=1e-9)
    assert slabs[:, 7].sum() / TWO32 == pytest.approx(m.potential, rel=1e-9)
    total = raw[bins * 8:] / TWO32
    assert total.sum() == pytest.approx(m.virial, rel=1e-9)
    np.testing.assert_allclose(slabs[:, 4:7].sum(axis=0) / TWO32, total, rtol=1e-9)
    p, _ = sim.read_particles()
    # diagonal, pair by pair, in the box as the simulation holds it (f32 sides)
    _, w_tensor = oracle.all_pairs(p, list(sim.params.box), RC, False)
    np.testing.assert_allclose(total, w_tensor, rtol=1e-9)


def test_profile_samples_accumulate_and_clear(oracle):
    pos, vel, box = sc_liquid(oracle, n_side=8)
    sim = oracle.Sim(oracle.jitter(pos, 6, 0.3), vel, [box] * 3)
    sim.set_profile_bins(9)
    sim.run(5)
    sim.sample_profile()
    one, _ = sim.read_profile()
    sim.sample_profile()
    two, samples = sim.read_profile()
    assert samples == 2
    np.testing.assert_array_equal(two, 2 * one)
    sim.clear_profile()
    zero, samples = sim.read_profile()
    assert samples == 0 and not zero.any()
