"""Checkpoint and exact resume: a handle loaded from a state file steps on
to the same bits as the handle that wrote it (NVE, Langevin with its
step-keyed noise, Nose-Hoover, heat exchange and the heat current,
profiles, a batch with bricks cut into parts, an all-NVE batch saved
between drift-ahead steps), and both agree with the oracle run without
interruption."""
import numpy as np
import pytest

from helpers import sc_liquid

pytestmark = pytest.mark.gpu


def same_state(a, b):
    pa, pb = a.read_particle_array(), b.read_particle_array()
    np.testing.assert_array_equal(pa["pos"], pb["pos"])
    np.testing.assert_array_equal(pa["vel"], pb["vel"])


def test_resume_nve_matches_the_oracle(oracle, lj, tmp_path):
    pos, vel, box = sc_liquid(oracle, n_side=10)
    st = lj.Settings([box] * 3, 2.5, 0.005)
    sim = lj.LennardJones((pos, vel), st)
    ref = oracle.Sim(pos, vel, [box] * 3)
    sim.run(137)
    ref.run(137)
    path = tmp_path / "nve.state"
    sim.save(path)
    back = lj.LennardJones.load(path)
    assert back.steps() == sim.steps() == 137
    for steps in (1, 90):
        sim.run(steps)
        back.run(steps)
        ref.run(steps)
        m, b, r = sim.measure_now(), back.measure_now(), ref.measure()
        assert list(m.raw) == list(b.raw) == list(r.raw)
        assert m.list_rebuilds == b.list_rebuilds
        same_state(sim, back)
    rp, rv = ref.read_particles()
    np.testing.assert_array_equal(back.read_particle_array()["pos"], rp)


def test_resume_langevin_and_nose_hoover(oracle, lj, tmp_path):
    pos, vel, box = sc_liquid(oracle, n_side=8)
    for k, th in enumerate([lj.Langevin(1.2, 1.0, 77), lj.NoseHoover(0.9, 0.5)]):
        st = lj.Settings([box] * 3, 2.5, 0.004, thermostat=th, shifted_force=True)
        sim = lj.LennardJones((pos, vel), st)
        sim.run(60)
        sim.set_temperature(1.4)
        sim.run(41)
        path = tmp_path / f"t{k}.state"
        sim.save(path)
        back = lj.LennardJones.load(path)
        assert back.target_temperature() == 1.4
        sim.run(120)
        back.run(120)
        m, b = sim.measure_now(), back.measure_now()
        assert list(m.raw) == list(b.raw)
        assert m.thermostat_energy == b.thermostat_energy
        same_state(sim, back)


def test_resume_heat_exchange_profiles_and_heat_current(oracle, lj, tmp_path):
    size = [9.0, 9.0, 24.0]
    count = int(round(0.8442 * np.prod(size)))
    pos, vel = oracle.lattice_box(count, size, 1.0, 11)
    st = lj.Settings(size, 2.5, 0.002, shifted_force=True, profile_bins=8)
    sim = lj.LennardJones((pos, vel), st)
    sim.run(50)
    sim.set_heat_exchange(lj.HeatExchange((0.0, 3.0), (12.0, 15.0), 0.02))
    sim.record_heat_current(5)
    sim.run(30)
    sim.sample_profile()
    sim.sample_heat_current()
    path = tmp_path / "heat.state"
    sim.save(path)
    back = lj.LennardJones.load(path)
    assert back.heat_flux() == sim.heat_flux()
    for s in (sim, back):
        s.run(40)
        s.sample_profile()
        s.sample_heat_current()
    assert list(sim.measure_now().raw) == list(back.measure_now().raw)
    np.testing.assert_array_equal(sim.read_profile_raw()[0], back.read_profile_raw()[0])
    assert sim.read_profile_raw()[1] == back.read_profile_raw()[1] == 2
    np.testing.assert_array_equal(sim.read_heat_current_raw(), back.read_heat_current_raw())
    same_state(sim, back)


def test_resume_batch_with_cut_bricks(oracle, lj, tmp_path):
    pos_a, vel_a, box_a = sc_liquid(oracle, n_side=14)
    pos_b, vel_b, box_b = sc_liquid(oracle, n_side=9)
    specs = [((pos_a, vel_a), lj.Settings([box_a] * 3, 2.5, 0.005, thermostat=lj.Langevin(0.9, 1.0, 5))),
             ((pos_b, vel_b), lj.Settings([box_b] * 3, 2.5, 0.005))]
    with lj.debug_plan(stage_fill=100.0):  # bricks that are cut into parts
        batch = lj.Batch(specs)
    batch.run(70)
    assert batch.debug_counters()["stage_overflow_builds"] > 0
    path = tmp_path / "batch.state"
    batch.save(path)
    back = lj.Batch.load(path)  # planned with the rules, it keeps the saved bricks
    assert back.sizes == batch.sizes and back.replicas == 2
    assert back.debug_counters()["bricks"] == batch.debug_counters()["bricks"]
    batch.run(150)
    back.run(150)
    for m, b in zip(batch.measure_now(), back.measure_now()):
        assert list(m.raw) == list(b.raw)
    a, b = batch.read_particle_array(), back.read_particle_array()
    np.testing.assert_array_equal(a["pos"], b["pos"])
    np.testing.assert_array_equal(a["vel"], b["vel"])
    with pytest.raises(lj.LennardJonesError, match="Batch.load"):
        lj.LennardJones.load(path)


def test_resume_nve_batch_with_cut_bricks(oracle, lj, tmp_path):
    # every replica NVE: saved between drift-ahead steps (the next positions
    # already formed in the second set), loaded, and both handles stepped
    # on; each replica also against the oracle run without a break
    parts = [sc_liquid(oracle, n_side=n) for n in (12, 9)]
    specs = [((pos, vel), lj.Settings([box] * 3, 2.5, 0.005)) for pos, vel, box in parts]
    with lj.debug_plan(stage_fill=100.0, parts=True):  # bricks that are cut into parts
        batch = lj.Batch(specs)
    batch.run(70)
    assert batch.debug_counters()["stage_overflow_builds"] > 0
    path = tmp_path / "nve_batch.state"
    batch.save(path)
    back = lj.Batch.load(path)
    for steps in (1, 149):
        batch.run(steps)
        back.run(steps)
    ms, bs = batch.measure_now(), back.measure_now()
    a, b = batch.read_particle_array(), back.read_particle_array()
    np.testing.assert_array_equal(a["pos"], b["pos"])
    np.testing.assert_array_equal(a["vel"], b["vel"])
    off = 0
    for (pos, vel, box), m, k in zip(parts, ms, bs):
        assert list(m.raw) == list(k.raw)
        ref = oracle.Sim(pos, vel, [box] * 3)
        ref.run(220)
        assert list(m.raw) == list(ref.measure().raw)
        rp, rv = ref.read_particles()
        n = len(pos)
        np.testing.assert_array_equal(a["pos"][off:off + n], rp)
        np.testing.assert_array_equal(a["vel"][off:off + n], rv)
        off += n
        ref.close()
