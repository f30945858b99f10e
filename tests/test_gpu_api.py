"""GPU tests of the lennard_jones API semantics (lennard_jones.rs:1092-1375),
Langevin trajectories and batched replicas, all against the oracle."""
import numpy as np
import pytest

from helpers import jittered_lattice, sc_liquid

pytestmark = pytest.mark.gpu


def settings(lj, box, **kw):
    return lj.Settings([box] * 3 if np.isscalar(box) else list(box), kw.pop("cutoff", 2.5), kw.pop("dt", 0.005), **kw)


def test_measurement_ring_and_poll(oracle, lj):
    # lennard_jones.rs:1141-1246: three read-back slots, poll oldest first
    pos, vel, box = sc_liquid(oracle, n_side=8)
    sim = lj.LennardJones((pos, vel), settings(lj, box))
    assert sim.measure()
    sim.step(5)
    assert sim.measure()
    sim.step(5)
    assert sim.measure()
    assert not sim.measure()  # ring full
    sim.after_submit()
    sim.run(0)
    got = sim.poll()
    assert [m.step for m in got] == [0, 5, 10]
    m = sim.measure_now()
    assert m.step == 10 and m.count == 512
    assert m.volume == pytest.approx(np.prod(sim.box_size()))
    assert sim.steps() == 10 and sim.time() == pytest.approx(0.05)


def test_measure_now_with_a_full_ring_raises(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=8)
    sim = lj.LennardJones((pos, vel), settings(lj, box))
    for _ in range(3):
        assert sim.measure()
    with pytest.raises(lj.LennardJonesError, match="already in flight") as e:
        sim.measure_now()
    assert e.value.code == 3
    sim.run(0)
    assert len(sim.poll()) == 3


def test_read_particles_round_trip(oracle, lj):
    # new -> read_particles gives the input back, positions wrapped into the box
    pos, vel = jittered_lattice(oracle)
    pos = pos + np.array([23.0, -11.0, 5.5])  # outside the box
    sim = lj.LennardJones((pos, vel), settings(lj, 11.0))
    out = sim.read_particle_array()
    L = sim.box_size()[0]
    wrapped = np.fmod(pos, L)
    wrapped[wrapped < 0] += L
    np.testing.assert_array_equal(out["pos"], wrapped)
    np.testing.assert_array_equal(out["vel"], vel)
    sim.run(30)  # after rebuilds and reordering: still the original order
    ref = oracle.Sim(pos, vel, [11.0] * 3)
    ref.run(30)
    rp, rv = ref.read_particles()
    out = sim.read_particle_array()
    np.testing.assert_array_equal(out["pos"], rp)
    np.testing.assert_array_equal(out["vel"], rv)
    assert np.all((out["pos"] >= 0) & (out["pos"] < L))


def test_thermostat_accessors_and_errors(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=8)
    nve = lj.LennardJones((pos, vel), settings(lj, box))
    assert nve.target_temperature() is None
    with pytest.raises(lj.LennardJonesError, match="set_temperature needs a thermostat"):
        nve.set_temperature(1.0)
    lv = lj.LennardJones((pos, vel), settings(lj, box, thermostat=lj.Langevin(1.5, 1.0, 3)))
    assert lv.target_temperature() == 1.5
    lv.set_temperature(2.0)
    assert lv.target_temperature() == 2.0
    with pytest.raises(lj.LennardJonesError, match="profile_bins"):
        lv.sample_profile()


def test_langevin_trajectory_bit_exact(oracle, lj):
    # counter-based noise: the GPU reproduces the oracle's BAOAB run bit for bit
    pos, vel, box = sc_liquid(oracle, n_side=9, temperature=1.2)
    sim = lj.LennardJones((pos, vel), settings(lj, box, thermostat=lj.Langevin(1.5, 0.5, 11), shifted_force=True))
    ref = oracle.Sim(pos, vel, [box] * 3, thermostat=1, temperature=1.5, friction=0.5, seed=11, shifted_force=True)
    sim.run(120)
    ref.run(120)
    m, r = sim.measure_now(), ref.measure()
    assert list(m.raw) == list(r.raw)
    p, v, a, o = sim.debug_state()
    rp, rv, ra, ro = ref.state()
    np.testing.assert_array_equal(o, ro)
    np.testing.assert_array_equal(p, rp)
    np.testing.assert_array_equal(v, rv)


def test_thermostatted_reruns_are_bit_identical(oracle, lj):
    # tests/thermodynamics.rs:295-326
    pos, vel = oracle.lattice(2000, 16.0, 1.5, 11)
    s = settings(lj, 16.0, dt=0.004, thermostat=lj.Langevin(1.5, 0.5, 3))

    def run():
        sim = lj.LennardJones((pos, vel), s)
        sim.run(300)
        m = sim.measure_now()
        return sim.read_particle_array().tobytes(), m.potential, m.virial

    a, b = run(), run()
    assert a[0] == b[0]
    assert (a[1], a[2]) == (b[1], b[2])


def test_batch_equals_single_runs(oracle, lj):
    # replicas of different size, density, thermostat and seed in one handle
    specs = [(6, 0.8442, None), (8, 0.6, lj.Langevin(1.3, 1.0, 5)), (7, 0.3, lj.Langevin(0.9, 0.5, 6)),
             (10, 0.8, None)]
    reps = []
    for n_side, rho, th in specs:
        n = n_side ** 3
        box = (n / rho) ** (1 / 3)
        pos, vel = oracle.lattice(n, box, 1.0, n_side)
        reps.append(((pos, vel), settings(lj, box, thermostat=th)))
    batch = lj.Batch(reps)
    batch.run(60)
    ms = batch.measure_now()
    allp = batch.read_particle_array()
    off = 0
    for (parts, st), m in zip(reps, ms):
        single = lj.LennardJones(parts, st)
        single.run(60)
        s = single.measure_now()
        assert list(m.raw) == list(s.raw)
        assert m.list_rebuilds == s.list_rebuilds
        n = single.count()
        got = allp[off:off + n]
        exp = single.read_particle_array()
        np.testing.assert_array_equal(got["pos"], exp["pos"])
        np.testing.assert_array_equal(got["vel"], exp["vel"])
        off += n


def test_dense_cluster_takes_the_global_memory_path(oracle, lj):
    # a dense blob in a dilute box: bricks around it stage more than fits in
    # LDS and read positions from global memory; results stay exact
    rng = np.random.default_rng(2)
    box = 40.0  # rc 3.5: 10 cells of 4.0 per axis
    blob, _ = oracle.lattice(8000, 15.0, 1.0, 4)  # spacing 0.75: about 150 atoms per filled cell
    blob = blob + 11.5
    gas = rng.uniform(0, box, (300, 3))
    d = gas[:, None, :] - blob[None, :, :]
    gas = gas[np.sqrt((d ** 2).sum(-1)).min(axis=1) > 1.0]
    pos = np.concatenate([blob, gas])
    vel = rng.standard_normal(pos.shape) * 0.5
    sim = lj.LennardJones((pos, vel), settings(lj, box, cutoff=3.5, dt=0.001))
    ref = oracle.Sim(pos, vel, [box] * 3, cutoff=3.5, dt=0.001)
    sim.run(20)
    ref.run(20)
    m, r = sim.measure_now(), ref.measure()
    assert list(m.raw) == list(r.raw)
    np.testing.assert_array_equal(sim.debug_state()[0], ref.state()[0])
    # some 3 x 3 x 3 neighbourhood holds more atoms than a brick can stage
    counts = np.diff(ref.cells()[1].astype(np.int64)).reshape(10, 10, 10)
    hood = sum(np.roll(counts, (a, b, c), axis=(0, 1, 2)) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1))
    assert hood.max() > 3200
    assert sim.debug_counters()["global_cell_builds"] > 0


@pytest.mark.parametrize("layout", ["dense", "sparse"])
def test_bricks_cut_into_parts_stay_exact(oracle, lj, layout):
    # bricks planned far beyond what LDS holds (8 x 8 home cells of a
    # liquid stage ~5700 atoms): each is cut into runs of home rows (dense
    # layout, 3200 atoms) or of home cells (sparse layout, 1024); lists,
    # forces, measurements, profiles and the heat current stay exact
    size = [24.0, 24.0, 12.0]
    count = int(round(0.8442 * np.prod(size)))
    pos, vel = oracle.lattice_box(count, size, 1.0, 5)
    st = lj.Settings(size, 2.5, 0.002, shifted_force=True, profile_bins=6)
    with lj.debug_plan(stage_fill=100.0, sparse_fill=100.0 if layout == "sparse" else -1.0, parts=True):
        sim = lj.LennardJones((pos, vel), st)
    ref = oracle.Sim(pos, vel, size, dt=0.002, shifted_force=True)
    ref.set_profile_bins(6)
    assert sim.debug_grid()["cells"] == [8, 8, 4]
    sim.record_heat_current(2)
    ref.record_heat_current(2)
    for steps in (1, 60):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert list(m.raw) == list(r.raw)
        np.testing.assert_array_equal(sim.debug_state()[0], ref.state()[0])
        sim.sample_profile()
        ref.sample_profile()
        sim.sample_heat_current()
        ref.sample_heat_current()
    np.testing.assert_array_equal(sim.read_profile_raw()[0], ref.read_profile()[0])
    np.testing.assert_array_equal(sim.read_heat_current_raw(), ref.read_heat_current())
    c = sim.debug_counters()
    assert c["bricks"] == 4 and c["stage_overflow_builds"] >= 4 * c["rebuilds"]
    assert c["global_cell_builds"] == 0
    assert c["sparse_bricks"] == (4 if layout == "sparse" else 0)


@pytest.mark.parametrize("layout", ["dense", "sparse"])
def test_outgrown_bricks_run_whole_without_parts(oracle, lj, layout):
    # the bricks of the test above, on a handle without the PARTS launches:
    # each outgrown brick runs whole, every position read through the cell
    # tables from global memory; lists, forces, measurements, profiles and
    # the heat current stay exact
    size = [24.0, 24.0, 12.0]
    count = int(round(0.8442 * np.prod(size)))
    pos, vel = oracle.lattice_box(count, size, 1.0, 5)
    st = lj.Settings(size, 2.5, 0.002, shifted_force=True, profile_bins=6)
    with lj.debug_plan(stage_fill=100.0, sparse_fill=100.0 if layout == "sparse" else -1.0, parts=False):
        sim = lj.LennardJones((pos, vel), st)
    ref = oracle.Sim(pos, vel, size, dt=0.002, shifted_force=True)
    ref.set_profile_bins(6)
    sim.record_heat_current(2)
    ref.record_heat_current(2)
    for steps in (1, 60):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert list(m.raw) == list(r.raw)
        np.testing.assert_array_equal(sim.debug_state()[0], ref.state()[0])
        sim.sample_profile()
        ref.sample_profile()
        sim.sample_heat_current()
        ref.sample_heat_current()
    np.testing.assert_array_equal(sim.read_profile_raw()[0], ref.read_profile()[0])
    np.testing.assert_array_equal(sim.read_heat_current_raw(), ref.read_heat_current())
    c = sim.debug_counters()
    assert c["bricks"] == 4 and c["stage_overflow_builds"] >= 4 * c["rebuilds"]
    assert c["global_cell_builds"] >= 4 * c["rebuilds"]
    assert c["sparse_bricks"] == (4 if layout == "sparse" else 0)


def test_gas_on_sparse_bricks_matches_the_oracle(oracle, lj):
    # a gas at rho 0.05: its layers take the sparse layout (128-thread
    # workgroups staging up to 1024 atoms); Langevin trajectory bit-exact
    side = float(np.float32((3000 / 0.05) ** (1.0 / 3.0)))
    parts = lj.random_insertion_array(3000, [side] * 3, 0.9, 1.5, 17)
    pos, vel = np.ascontiguousarray(parts["pos"]), np.ascontiguousarray(parts["vel"])
    sim = lj.LennardJones((pos, vel), settings(lj, side, thermostat=lj.Langevin(1.5, 1.0, 23)))
    ref = oracle.Sim(pos, vel, [side] * 3, dt=0.005, thermostat=1, temperature=1.5, friction=1.0, seed=23)
    c = sim.debug_counters()
    assert c["sparse_bricks"] == c["bricks"] > 0
    for steps in (1, 150):
        sim.run(steps)
        ref.run(steps)
        assert list(sim.measure_now().raw) == list(ref.measure().raw)
        np.testing.assert_array_equal(sim.debug_state()[0], ref.state()[0])


def test_two_cells_per_axis(oracle, lj):
    # 216 atoms at rho 0.8442: L = 6.43, two cells of 3.2 per axis
    pos, vel, box = sc_liquid(oracle, n_side=6)
    sim = lj.LennardJones((pos, vel), settings(lj, box))
    ref = oracle.Sim(pos, vel, [box] * 3)
    assert sim.debug_grid()["cells"] == [2, 2, 2]
    m, r = sim.measure_now(), ref.measure()
    L = sim.box_size()[0]
    pe, w = oracle.all_pairs(np.mod(pos, L), [L] * 3, 2.5, False)
    assert abs((m.potential - pe) / pe) < 1e-12
    sim.run(50)
    ref.run(50)
    assert list(sim.measure_now().raw) == list(ref.measure().raw)


def test_empty_and_tiny_systems(lj):
    sim = lj.LennardJones((np.zeros((0, 3)), np.zeros((0, 3))), settings(lj, 10.0))
    sim.run(10)
    m = sim.measure_now()
    assert m.count == 0 and m.kinetic == 0.0 and m.potential == 0.0
    assert sim.read_particles() == []
    one = lj.LennardJones(([[1.0, 2.0, 3.0]], [[0.1, 0.0, 0.0]]), settings(lj, 10.0))
    one.run(100)
    m = one.measure_now()
    assert m.potential == 0.0
    p = one.read_particles()[0]
    assert p.pos[0] == pytest.approx(1.0 + 100 * 0.005 * 0.1)


def test_fp64_fma_rate_is_plausible(lj):
    # the FMA-chain measurement that the force kernel's VALU roofline sits
    # beside: positive and not above MI355X's published 78.6 TFLOP/s by
    # more than timing noise
    tf = lj.fp64_fma_rate(3)
    assert 20.0 < tf < 1.2 * 78.6


def test_fp32_form_is_the_same_on_cut_bricks(oracle, lj):
    # the FP32 form sums each list octet in FP32 by list position, so bricks
    # cut into parts (other staging, same list order) give the same bits
    size = [24.0, 24.0, 12.0]
    count = int(round(0.8442 * np.prod(size)))
    pos, vel = oracle.lattice_box(count, size, 1.0, 5)
    runs = []
    for fill in (0.0, 100.0):
        with lj.debug_plan(stage_fill=fill):
            sim = lj.LennardJones((pos, vel), lj.Settings(size, 2.5, 0.002, precision=1))
        sim.run(40)
        runs.append((list(sim.measure_now().raw), sim.debug_state()[0], sim.debug_counters()))
    assert runs[0][2]["stage_overflow_builds"] == 0 and runs[1][2]["stage_overflow_builds"] > 0
    assert runs[0][0] == runs[1][0]
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
