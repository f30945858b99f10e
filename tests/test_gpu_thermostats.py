"""Nose-Hoover on the GPU (lennard_jones.wgsl:309-397, lennard_jones.rs:1107-1127):
bit-exact trajectories and extended energy against the oracle, retargeting,
batches, and the reference's canonical-ensemble check
(tests/thermodynamics.rs:329-381)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def nh_pair(oracle, lj, n=1000, rho=0.3, T=1.5, start_T=0.5, period=0.5, dt=0.002, seed=21):
    box = (n / rho) ** (1 / 3)
    pos, vel = oracle.lattice(n, box, start_T, seed)
    st = lj.Settings([box] * 3, 2.5, dt, thermostat=lj.NoseHoover(T, period), shifted_force=True)
    sim = lj.LennardJones((pos, vel), st)
    ref = oracle.Sim(pos, vel, [box] * 3, dt=dt, shifted_force=True, thermostat=2, temperature=T, period=period)
    return sim, ref, st, (pos, vel)


def test_nose_hoover_trajectory_is_bit_exact(oracle, lj):
    sim, ref, _, _ = nh_pair(oracle, lj)
    for steps in (1, 37, 150):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert list(m.raw) == list(r.raw)
        assert m.thermostat_energy == r.thermostat_energy
        p, v = sim.read_particle_array()["pos"], sim.read_particle_array()["vel"]
        rp, rv = ref.read_particles()
        np.testing.assert_array_equal(p, rp)
        np.testing.assert_array_equal(v, rv)
    assert m.thermostat_energy != 0.0


def test_nose_hoover_retarget_matches_the_oracle(oracle, lj):
    sim, ref, _, _ = nh_pair(oracle, lj, n=512)
    sim.run(100)
    ref.run(100)
    sim.set_temperature(0.9)
    ref.set_temperature(0.9)
    assert sim.target_temperature() == 0.9
    sim.run(100)
    ref.run(100)
    m, r = sim.measure_now(), ref.measure()
    assert list(m.raw) == list(r.raw)
    assert m.thermostat_energy == r.thermostat_energy


def test_nose_hoover_in_a_batch_equals_single_runs(oracle, lj):
    _, _, st, parts = nh_pair(oracle, lj, n=729)
    box = (512 / 0.5) ** (1 / 3)
    pos2, vel2 = oracle.lattice(512, box, 1.0, 3)
    specs = [(parts, st), ((pos2, vel2), lj.Settings([box] * 3, 2.5, 0.002, thermostat=lj.Langevin(1.0, 1.0, 4),
                                                     shifted_force=True))]
    batch = lj.Batch(specs)
    batch.run(120)
    ms = batch.measure_now()
    for (p, s), m in zip(specs, ms):
        single = lj.LennardJones(p, s)
        single.run(120)
        x = single.measure_now()
        assert list(m.raw) == list(x.raw)
        assert m.thermostat_energy == x.thermostat_energy


def test_nose_hoover_samples_the_canonical_ensemble(oracle, lj):
    # tests/thermodynamics.rs:329-381: N 1000 at rho 0.3, T 1.5, dt 0.002,
    # period 0.5, force-shifted; start cold at T 0.5; after 10,000 steps,
    # 2000 samples 20 steps apart
    sim, _, _, _ = nh_pair(oracle, lj)
    n = 1000
    sim.run(10_000)
    dof = 3.0 * n - 3.0
    start = sim.measure_now()
    temps = []
    for _ in range(2000):
        sim.run(20)
        temps.append(sim.measure_now().temperature(dof))
    end = sim.measure_now()
    temps = np.array(temps)
    mean = temps.mean()
    ratio = temps.var(ddof=1) / mean ** 2
    canonical = 2.0 / dof
    drift = (end.extended_energy() - start.extended_energy()) / abs(start.extended_energy())
    print(f"Nose-Hoover: mean T {mean:.4f}; Var(T)/T^2 {ratio:.3e} vs 2/g {canonical:.3e} "
          f"({(ratio / canonical - 1) * 100:+.1f}%); extended energy drift {drift:.2e}")
    assert abs(mean / 1.5 - 1) < 0.01
    assert abs(ratio / canonical - 1) < 0.15
    assert abs(drift) < 1e-4
    assert end.list_overflow == 0
