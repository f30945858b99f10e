"""The brick planner's tuner and re-planning on the GPU.

Bricks decide only how neighbourhoods are staged in LDS, never which pairs
a list holds or in what order, so a tuned or re-planned run stays bit for
bit on the oracle's trajectory. The lists rebuilt for a new plan come from
the positions of the last list build, so a plan can change at any step.
"""
import numpy as np
import pytest

from helpers import sc_liquid

pytestmark = pytest.mark.gpu


def pair(oracle, lj, pos, vel, size, thermostat=None):
    kind, temp, fric, seed = 0, 0.0, 0.0, 1
    if thermostat is not None:
        kind, temp, fric, seed = 1, thermostat.temperature, thermostat.friction, thermostat.seed
    ref = oracle.Sim(pos, vel, size, dt=0.005, thermostat=kind, temperature=temp, friction=fric, seed=seed)
    return ref


def test_tuning_mid_run_keeps_the_trajectory(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=26)  # 17,576 atoms
    st = lj.Settings([box] * 3, 2.5, 0.005, skin=0.3)
    sim = lj.LennardJones((pos, vel), st)
    ref = pair(oracle, lj, pos, vel, [box] * 3)
    sim.run(37)  # between list builds
    taken = sim.tune(5)
    info = sim.plan_info()
    assert taken == 15 and len(info["trials"]) == 3
    assert [t["fill"] for t in info["trials"]] == [0.8, 0.9, 1.0]
    assert info["fill"] in (0.8, 0.9, 1.0)
    assert all(t["force_us"] > 0 and t["build_us"] > 0 for t in info["trials"])
    sim.run(43)
    ref.run(37 + taken + 43)
    m, r = sim.measure_now(), ref.measure()
    assert list(m.raw) == list(r.raw)
    np.testing.assert_array_equal(sim.debug_state()[0], ref.state()[0])
    np.testing.assert_array_equal(sim.debug_state()[1], ref.state()[1])


def test_large_handle_tunes_itself(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=42)  # 74,088 atoms
    sim = lj.LennardJones((pos, vel), lj.Settings([box] * 3, 2.5, 0.005, skin=0.3))
    sim.run(10)  # too few to tune
    assert sim.plan_info()["trials"] == []
    sim.run(100)  # tunes on the way: 3 x 6 of these steps
    assert len(sim.plan_info()["trials"]) == 3
    assert sim.steps() == 110
    ref = pair(oracle, lj, pos, vel, [box] * 3)
    ref.run(110)
    assert list(sim.measure_now().raw) == list(ref.measure().raw)


def test_cut_bricks_are_planned_again(oracle, lj):
    # planned far too large (every brick cut into parts); once the plan is
    # no longer fixed, the periodic check plans the replica again from its
    # occupancy, its bricks stop being cut, and the trajectory is the
    # oracle's throughout
    size = [24.0, 24.0, 12.0]
    count = int(round(0.8442 * np.prod(size)))
    pos, vel = oracle.lattice_box(count, size, 1.0, 5)
    thermo = lj.Langevin(1.0, 1.0, 77)
    with lj.debug_plan(stage_fill=100.0):
        sim = lj.LennardJones((pos, vel), lj.Settings(size, 2.5, 0.005, skin=0.3, thermostat=thermo))
    ref = pair(oracle, lj, pos, vel, size, thermostat=thermo)
    sim.run(600)
    c0 = sim.debug_counters()
    assert c0["stage_overflow_builds"] > 0
    sim.run(600)  # a check (every 256 steps) arrives and plans again
    info = sim.plan_info()
    c1 = sim.debug_counters()
    assert info["replans"] >= 1
    assert c1["bricks"] > c0["bricks"]
    sim.run(400)
    c2 = sim.debug_counters()
    assert c2["stage_overflow_builds"] == c1["stage_overflow_builds"]  # no more cuts
    ref.run(1600)
    m, r = sim.measure_now(), ref.measure()
    assert list(m.raw) == list(r.raw)
    np.testing.assert_array_equal(sim.debug_state()[0], ref.state()[0])


def test_batch_tuned_mid_run_equals_single_runs(lj):
    rng = np.random.default_rng(21)
    specs = []
    for k in range(6):
        n = int(rng.integers(2000, 9000))
        rho = float(rng.uniform(0.05, 0.85))
        temp = float(rng.uniform(0.8, 1.6))
        if rho >= 0.3:
            parts, side = lj.fcc_count_array(n, rho, temp, 40 + k)
        else:
            side = float(np.float32((n / rho) ** (1.0 / 3.0)))
            parts = lj.random_insertion_array(n, [side] * 3, 0.9, temp, 40 + k)
        specs.append((parts, lj.Settings([side] * 3, 2.5, 0.005, skin=0.3,
                                         thermostat=lj.Langevin(temp, 1.0, 90 + k))))
    batch = lj.Batch(specs)
    batch.run(25)
    taken = batch.tune(4)
    batch.run(31)
    ms = batch.measure_now()
    for (parts, st), m in zip(specs, ms):
        single = lj.LennardJones(parts, st)
        single.run(25 + taken + 31)
        assert list(single.measure_now().raw) == list(m.raw)
        single.close()
