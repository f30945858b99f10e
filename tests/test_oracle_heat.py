"""The oracle's heat current (lennard_jones.wgsl:872-947) against the f64
all-pairs sum of the reference's tests/heat.rs:54-90, and its heat
exchange (eHEX/a, lennard_jones.wgsl:732-870) against closed forms and the
bookkeeping of tests/heat.rs:158-212."""
import numpy as np
import pytest

TWO32 = 2.0 ** 32
RC = 2.5


def sf_pair(r2):
    """Shifted-force (U, r.F) in f64 (tests/heat.rs:39-50)."""
    def lj(r2):
        s6 = 1.0 / (r2 * r2 * r2)
        return 4.0 * (s6 * s6 - s6), 24.0 * (2.0 * s6 * s6 - s6)
    u, rf = lj(r2)
    uc, rfc = lj(RC * RC)
    r = np.sqrt(r2)
    fc = rfc / RC
    return u - uc + (r - RC) * fc, rf - r * fc


def heat_current_all_pairs(pos, vel, size):
    """J = sum_i e_i v_i + 1/2 sum_ij (F_ij . v_i) r_ij over all pairs,
    nearest image (tests/heat.rs:54-90), vectorised."""
    size = np.asarray(size)
    d = pos[None, :, :] - pos[:, None, :]          # d[i, j] = r_j - r_i
    d -= size * np.round(d / size)
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, 1e9)
    inside = r2 < RC * RC
    u, rf = sf_pair(np.where(inside, r2, 1.0))
    u = np.where(inside, u, 0.0)
    rf = np.where(inside, rf, 0.0)
    e = 0.5 * (vel * vel).sum(-1) + 0.5 * u.sum(1)
    dv = (d * vel[:, None, :]).sum(-1)            # d_ij . v_i
    pairs = (d * (dv * rf / r2)[:, :, None]).sum(1)
    return (e[:, None] * vel + 0.5 * pairs).sum(0)


def liquid(oracle, size, t, seed=11):
    count = int(round(0.8442 * np.prod(size)))
    return oracle.lattice_box(count, size, t, seed)


def test_heat_current_matches_all_pairs(oracle):
    size = [9.0, 9.0, 12.0]
    pos, vel = liquid(oracle, size, 1.0)
    sim = oracle.Sim(pos, vel, size, dt=0.002, shifted_force=True)
    sim.run(300)  # melt the lattice first
    sim.record_heat_current(3)
    sim.sample_heat_current()
    sim.run(5)
    sim.sample_heat_current()
    raw = sim.read_heat_current()
    assert raw.shape == (2, 3)
    p, v = sim.read_particles()
    want = heat_current_all_pairs(p, v, list(sim.params.box))
    scale = np.sqrt(len(p)) * 10.0  # terms of order 1 with both signs (tests/heat.rs:124-127)
    np.testing.assert_allclose(raw[1] / TWO32, want, rtol=0, atol=1e-9 * scale)
    sim.sample_heat_current()
    with pytest.raises(ValueError):
        sim.sample_heat_current()


def test_heat_exchange_closed_form_without_forces(oracle):
    # a gas with no pair inside the cutoff: one step moves exactly `energy`
    # into the hot slab's kinetic energy about its centre of mass and out of
    # the cold one's, momentum of each slab unchanged
    rng = np.random.default_rng(4)
    size = [30.0, 30.0, 60.0]
    g = np.stack(np.meshgrid(np.arange(5), np.arange(5), np.arange(10), indexing="ij"), -1).reshape(-1, 3)
    pos = (g + 0.5) * np.array([6.0, 6.0, 6.0])
    vel = rng.normal(0, 1, pos.shape)
    dt, energy = 0.002, 0.05
    sim = oracle.Sim(pos, vel, size, dt=dt, shifted_force=True)
    sim.set_heat_exchange((0.0, 12.0), (30.0, 42.0), energy)

    def slab_state(p, v, lo, hi):
        z = np.mod(p[:, 2], size[2])
        m = (z >= lo) & (z < hi)
        vv = v[m]
        mom = vv.sum(0)
        return m, 0.5 * (vv * vv).sum() - 0.5 * (mom @ mom) / m.sum(), mom

    p0, v0 = sim.read_particles()
    hot0 = slab_state(p0 + v0 * dt, v0, 0.0, 12.0)
    cold0 = slab_state(p0 + v0 * dt, v0, 30.0, 42.0)
    sim.run(1)
    p1, v1 = sim.read_particles()
    hot1 = slab_state(p1, v1, 0.0, 12.0)
    cold1 = slab_state(p1, v1, 30.0, 42.0)
    assert hot0[0].sum() > 0 and cold0[0].sum() > 0
    assert hot1[1] - hot0[1] == pytest.approx(energy, rel=1e-9)
    assert cold1[1] - cold0[1] == pytest.approx(-energy, rel=1e-9)
    np.testing.assert_allclose(hot1[2], hot0[2], atol=1e-12)
    np.testing.assert_allclose(cold1[2], cold0[2], atol=1e-12)
    # atoms outside the slabs only drift
    outside = ~(hot0[0] | cold0[0])
    np.testing.assert_array_equal(v1[outside], v0[outside])
    with pytest.raises(ValueError):
        t = oracle.Sim(pos, vel, size, thermostat=1, temperature=1.0, friction=1.0)
        t.set_heat_exchange((0.0, 1.0), (2.0, 3.0), 0.1)


def test_heat_exchange_moves_energy_and_keeps_momentum(oracle):
    # tests/heat.rs:158-212 at a CPU-sized length: energy stays constant
    # (the drift slope is a small part of the energy moved), momentum is
    # kept, and the hot slab ends up warmer than the cold one
    size = [9.0, 9.0, 36.0]
    pos, vel = liquid(oracle, size, 1.0)
    sim = oracle.Sim(pos, vel, size, dt=0.002, shifted_force=True)
    sim.set_profile_bins(12)
    sim.run(1000)
    before = sim.measure()
    energy = 0.01
    sim.set_heat_exchange((0.0, 3.0), (18.0, 21.0), energy)
    time, total = [], []
    for _ in range(100):  # 20 time units: Verlet's own fluctuation (~0.05) hides a drift in fewer
        sim.run(100)
        sim.sample_profile()
        m = sim.measure()
        time.append(m.step * 0.002)
        total.append(m.kinetic + m.potential)
    after = sim.measure()
    rate = np.polyfit(time, total, 1)[0]
    duration = time[-1] - time[0]
    moved = energy / 0.002 * duration
    assert abs(rate * duration) < 1e-3 * moved
    for k in range(3):
        assert abs(after.momentum[k] - before.momentum[k]) < 1e-6 * after.count
    raw, samples = sim.read_profile()
    slabs = raw[: 12 * 8].reshape(12, 8) / TWO32
    temp = slabs[:, 1:4].sum(1) / (3 * slabs[:, 0])
    assert temp[0] > temp[6] + 0.04
