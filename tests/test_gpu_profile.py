"""GPU profiles along z (lennard_jones.wgsl:628-730) against the oracle.

Every (pair, slab, axis) share goes to 64-bit fixed point on its own in both,
so the raw sums are compared bit for bit: the LDS path (up to 256 slabs),
the global-atomic path (more slabs), shifted forces, a box of two cells per
axis, the global-memory staging fallback, and batched replicas.
"""
import numpy as np
import pytest

from helpers import sc_liquid

pytestmark = pytest.mark.gpu

TWO32 = 2.0 ** 32


def pair(oracle, lj, pos, vel, box, bins, cutoff=2.5, dt=0.005, shifted_force=False, precision=0):
    box3 = [box] * 3 if np.isscalar(box) else list(box)
    st = lj.Settings(box3, cutoff, dt, shifted_force=shifted_force, profile_bins=bins, precision=precision)
    sim = lj.LennardJones((pos, vel), st)
    ref = oracle.Sim(pos, vel, box3, cutoff=cutoff, dt=dt, shifted_force=shifted_force, precision=precision)
    ref.set_profile_bins(bins)
    return sim, ref


@pytest.mark.parametrize("bins", [1, 7, 40, 200, 256, 257, 300])
def test_profile_sums_bit_exact(oracle, lj, bins):
    print(f"profile parity, bins {bins}", flush=True)
    pos, vel, box = sc_liquid(oracle, n_side=12)
    pos = oracle.jitter(pos, 17, 0.25)
    sim, ref = pair(oracle, lj, pos, vel, box, bins)
    for steps in (25, 10):
        sim.run(steps)
        ref.run(steps)
        sim.sample_profile()
        ref.sample_profile()
    raw, samples = sim.read_profile_raw()
    r_raw, r_samples = ref.read_profile()
    assert samples == r_samples == 2
    np.testing.assert_array_equal(raw, r_raw)
    assert raw[0:bins * 8:8].sum() == 2 * len(pos) * 2 ** 32


def test_profile_shifted_force_and_averages(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=10)
    sim, ref = pair(oracle, lj, oracle.jitter(pos, 3, 0.2), vel, box, 16, shifted_force=True)
    for _ in range(3):
        sim.run(7)
        ref.run(7)
        sim.sample_profile()
        ref.sample_profile()
    raw, samples = sim.read_profile_raw()
    r_raw, _ = ref.read_profile()
    np.testing.assert_array_equal(raw, r_raw)
    p = sim.read_profile()
    assert p.samples == samples == 3
    sums = raw / TWO32 / 3
    assert [s.particles for s in p.slabs] == list(sums[0:16 * 8:8])
    assert p.virial == list(sums[16 * 8:])
    assert sum(s.particles for s in p.slabs) == pytest.approx(len(pos))


def test_profile_clear_and_errors(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=8)
    sim, ref = pair(oracle, lj, pos, vel, box, 5)
    sim.run(3)
    sim.sample_profile()
    sim.clear_profile()
    raw, samples = sim.read_profile_raw()
    assert samples == 0 and not raw.any()
    assert all(np.isnan(s.particles) for s in sim.read_profile().slabs)
    sim.sample_profile()
    ref.run(3)
    ref.sample_profile()
    np.testing.assert_array_equal(sim.read_profile_raw()[0], ref.read_profile()[0])
    plain = lj.LennardJones((pos, vel), lj.Settings([box] * 3, 2.5, 0.005))
    with pytest.raises(lj.LennardJonesError, match="profiles need Settings::profile_bins above zero"):
        plain.sample_profile()
    with pytest.raises(lj.LennardJonesError, match="profiles need Settings::profile_bins above zero"):
        plain.read_profile()
    plain.clear_profile()  # nothing to clear


def test_profile_two_cells_per_axis(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=6)
    sim, ref = pair(oracle, lj, pos, vel, box, 11)
    assert sim.debug_grid()["cells"] == [2, 2, 2]
    sim.run(40)
    ref.run(40)
    sim.sample_profile()
    ref.sample_profile()
    np.testing.assert_array_equal(sim.read_profile_raw()[0], ref.read_profile()[0])


def test_profile_global_memory_staging(oracle, lj):
    # a dense blob whose bricks cannot stage their neighbourhood in LDS
    rng = np.random.default_rng(2)
    box = 40.0
    blob, _ = oracle.lattice(8000, 15.0, 1.0, 4)
    blob = blob + 11.5
    gas = rng.uniform(0, box, (300, 3))
    d = gas[:, None, :] - blob[None, :, :]
    gas = gas[np.sqrt((d ** 2).sum(-1)).min(axis=1) > 1.0]
    pos = np.concatenate([blob, gas])
    vel = rng.standard_normal(pos.shape) * 0.5
    sim, ref = pair(oracle, lj, pos, vel, box, 64, cutoff=3.5, dt=0.001)
    sim.run(10)
    ref.run(10)
    sim.sample_profile()
    ref.sample_profile()
    np.testing.assert_array_equal(sim.read_profile_raw()[0], ref.read_profile()[0])
    assert sim.debug_counters()["global_cell_builds"] > 0


def test_batch_profiles_match_single_runs(oracle, lj):
    pos_a, vel_a, box_a = sc_liquid(oracle, n_side=8)
    pos_b, vel_b, box_b = sc_liquid(oracle, n_side=10)
    pos_b = oracle.jitter(pos_b, 8, 0.2)
    specs = [((pos_a, vel_a), lj.Settings([box_a] * 3, 2.5, 0.005, profile_bins=9)),
             ((pos_b, vel_b), lj.Settings([box_b] * 3, 2.5, 0.005)),
             ((pos_b, vel_b), lj.Settings([box_b] * 3, 2.5, 0.005, profile_bins=200))]
    batch = lj.Batch(specs)
    batch.run(12)
    batch.sample_profile()
    for r in (0, 2):
        single = lj.LennardJones(*specs[r])
        single.run(12)
        single.sample_profile()
        raw, samples = batch.read_profile_raw(r)
        assert samples == 1
        np.testing.assert_array_equal(raw, single.read_profile_raw()[0])
    with pytest.raises(lj.LennardJonesError, match="profiles need Settings::profile_bins above zero"):
        batch.read_profile_raw(1)


def test_fp32_forces_profile_adds_up(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=10)
    st = lj.Settings([box] * 3, 2.5, 0.005, profile_bins=20, precision=1)
    sim = lj.LennardJones((oracle.jitter(pos, 4, 0.2), vel), st)
    sim.run(20)
    sim.sample_profile()
    m = sim.measure_now()
    raw, _ = sim.read_profile_raw()
    slabs = raw[: 20 * 8].reshape(20, 8) / TWO32
    assert slabs[:, 0].sum() == len(pos)
    assert slabs[:, 1:4].sum() == pytest.approx(2.0 * m.kinetic, rel=1e-9)
    assert slabs[:, 7].sum() == pytest.approx(m.potential, rel=1e-5)
    assert (raw[20 * 8:] / TWO32).sum() == pytest.approx(m.virial, rel=1e-5)
