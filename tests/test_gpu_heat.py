"""Heat exchange (eHEX/a) and the heat current on the GPU against the oracle
bit for bit, and the reference's bookkeeping check (tests/heat.rs:158-212)
at its full length."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWO32 = 2.0 ** 32


def liquid_pair(oracle, lj, size, t=1.0, bins=0):
    count = int(round(0.8442 * np.prod(size)))
    pos, vel = oracle.lattice_box(count, size, t, 11)
    sim = lj.LennardJones((pos, vel), lj.Settings(size, 2.5, 0.002, shifted_force=True, profile_bins=bins))
    ref = oracle.Sim(pos, vel, size, dt=0.002, shifted_force=True)
    return sim, ref


def test_heat_current_series_is_bit_exact(oracle, lj):
    sim, ref = liquid_pair(oracle, lj, [9.0, 9.0, 12.0])
    sim.run(300)
    ref.run(300)
    sim.record_heat_current(4)
    ref.record_heat_current(4)
    for steps in (0, 7, 30):
        sim.run(steps)
        ref.run(steps)
        sim.sample_heat_current()
        ref.sample_heat_current()
    got = sim.read_heat_current_raw()
    np.testing.assert_array_equal(got, ref.read_heat_current())
    assert got.shape == (3, 3)
    np.testing.assert_array_equal(np.array(sim.read_heat_current()), got / TWO32)
    sim.sample_heat_current()
    with pytest.raises(lj.LennardJonesError, match="heat current series full"):
        sim.sample_heat_current()
    sim.record_heat_current(2)  # starts afresh
    assert sim.read_heat_current() == []


def test_heat_exchange_trajectory_is_bit_exact(oracle, lj):
    size = [9.0, 9.0, 24.0]
    sim, ref = liquid_pair(oracle, lj, size)
    sim.run(200)
    ref.run(200)
    sim.set_heat_exchange(lj.HeatExchange((0.0, 3.0), (12.0, 15.0), 0.02))
    ref.set_heat_exchange((0.0, 3.0), (12.0, 15.0), 0.02)
    assert sim.heat_flux() == pytest.approx(0.02 / (2 * 81.0 * 0.002))
    for steps in (1, 60, 240):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert list(m.raw) == list(r.raw)
        a = sim.read_particle_array()
        rp, rv = ref.read_particles()
        np.testing.assert_array_equal(a["pos"], rp)
        np.testing.assert_array_equal(a["vel"], rv)
    sim.set_heat_exchange(None)
    ref.set_heat_exchange(None, None, 0.0)
    assert sim.heat_flux() is None
    sim.run(50)
    ref.run(50)
    assert list(sim.measure_now().raw) == list(ref.measure().raw)


def test_heat_exchange_needs_constant_energy(lj):
    pos, vel = np.random.default_rng(1).uniform(0, 12, (100, 3)), np.zeros((100, 3))
    sim = lj.LennardJones((pos, vel), lj.Settings([12.0] * 3, 2.5, 0.002, thermostat=lj.Langevin(1.0, 1.0, 2)))
    with pytest.raises(lj.LennardJonesError, match="heat exchange runs at constant energy, without a thermostat"):
        sim.set_heat_exchange(lj.HeatExchange((0.0, 1.0), (6.0, 7.0), 0.01))
    with pytest.raises(lj.LennardJonesError, match="record_heat_current first"):
        sim.sample_heat_current()


def test_heat_exchange_moves_energy_and_keeps_momentum(oracle, lj):
    # tests/heat.rs:158-212: 9 x 9 x 36 liquid, 2000 steps, then 0.01 per
    # step from the slab 18..21 to 0..3 for 200 x 100 steps
    size = [9.0, 9.0, 36.0]
    sim, _ = liquid_pair(oracle, lj, size, bins=12)
    sim.run(2000)
    before = sim.measure_now()
    energy = 0.01
    sim.set_heat_exchange(lj.HeatExchange((0.0, 3.0), (18.0, 21.0), energy))
    time, total = [], []
    for _ in range(200):
        sim.run(100)
        sim.sample_profile()
        m = sim.measure_now()
        time.append(m.time)
        total.append(m.energy())
    after = sim.measure_now()
    rate = np.polyfit(time, total, 1)[0]
    duration = time[-1] - time[0]
    moved = energy / 0.002 * duration
    t = sim.read_profile().temperature()
    print(f"moved {moved:.0f}; drift {rate:+.2e} per time unit ({abs(rate * duration / moved):.1e} of moved); "
          f"slab temperatures {np.round(t, 3)}")
    assert abs(rate * duration) < 1e-3 * moved
    for k in range(3):
        assert abs(after.momentum[k] - before.momentum[k]) < 1e-6 * after.count
    assert t[0] > t[6] + 0.04
