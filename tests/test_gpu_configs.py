"""Every benchmark configuration through the HIP library against the oracle.

  SC-1000     N = 1000 SC, NVE, 1000 steps: test_gpu_parity.py::test_nve_trajectory_bit_exact
  FCC-1M      FCC 4 x 64^3 = 1,048,576, FP64 and FP32 forms: 20 steps bit-exact
              (cells, lists, fixed-point totals, positions) through list rebuilds
  FCC-8M      FCC 4 x 128^3 = 8,388,608: the oracle's step-0 totals and forces on
              the melted state in both forms; the FP32 form's force, PE and virial
              error against the oracle's FP64; NVE drift of both forms over 10^4 steps
  REPLICAS    the 512-replica Langevin batch (seed 7): every replica's step-0
              totals, seven chosen replicas over 100 steps, and the replica whose
              bricks were cut most often over all 5000 steps, against the oracle;
              no list overflow anywhere
  SLAB        one seed of the 2M-atom liquid-vapour slab (FP32 forces, force
              shifted, Langevin, 200 profile slabs) bit-exact over 20 steps with a
              profile sample; the 8-seed batch's profiles add up to its totals

The oracle runs on the host's cores (OpenMP); the largest case holds about
10 GB of host memory.
"""
import numpy as np
import pytest

from virial_particles import workloads as W

pytestmark = pytest.mark.gpu


def oracle_of(oracle, parts, st):
    """oracle.Sim with the same particles and settings."""
    t = st.thermostat
    kind, temp, fric, seed = 0, 0.0, 0.0, 1
    if t is not None:
        kind, temp, fric, seed = 1, t.temperature, t.friction, t.seed
    ref = oracle.Sim(np.ascontiguousarray(parts["pos"]), np.ascontiguousarray(parts["vel"]), list(st.box_size),
                     cutoff=st.cutoff, skin=st.skin, dt=st.dt, shifted_force=st.shifted_force,
                     precision=st.precision, thermostat=kind, temperature=temp, friction=fric, seed=seed)
    if st.profile_bins:
        ref.set_profile_bins(st.profile_bins)
    return ref


def same_lists(sim, ref):
    counts, entries = sim.debug_lists()
    r_counts, r_entries = ref.lists()
    np.testing.assert_array_equal(counts, r_counts)
    k = np.arange(entries.shape[1], dtype=np.uint32)
    live = k[None, :] < counts[:, None]
    assert np.array_equal(np.where(live, entries, 0), np.where(live, r_entries, 0))


def same_cells(sim, ref):
    for a, b in zip(sim.debug_cells(), ref.cells()):
        np.testing.assert_array_equal(a, b)


def same_state(sim, ref):
    p, v, a, o = sim.debug_state()
    rp, rv, ra, ro = ref.state()
    np.testing.assert_array_equal(o, ro)
    np.testing.assert_array_equal(p, rp)
    np.testing.assert_array_equal(v, rv)
    np.testing.assert_array_equal(a, ra)


@pytest.mark.parametrize("precision", [0, 1])
def test_fcc_1m_bit_exact_through_rebuilds(oracle, lj, precision):
    parts, st = W.fcc_liquid(64, precision=precision)
    sim = lj.LennardJones(parts, st)
    ref = oracle_of(oracle, parts, st)
    assert sim.count() == 1_048_576
    assert sim.debug_grid()["cells"] == [38, 38, 38] == list(ref.params.cells)
    assert sim.debug_grid()["capacity"] == ref.params.capacity
    assert list(sim.measure_now().raw) == list(ref.measure().raw)
    same_cells(sim, ref)
    same_lists(sim, ref)
    for steps in (9, 11):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert list(m.raw) == list(r.raw), f"after {sim.steps()} steps"
        assert m.list_rebuilds == r.list_rebuilds
    assert m.list_rebuilds >= 2  # the first lists and at least one rebuild
    same_state(sim, ref)
    same_cells(sim, ref)
    same_lists(sim, ref)


def test_fcc_8m_both_forms_against_the_oracle(oracle, lj):
    # melt the lattice (every force on the perfect lattice is zero by
    # symmetry), then both forms from the same state
    parts, st = W.fcc_liquid(128)
    melt = lj.LennardJones(parts, st)
    melt.run(1000)
    state = melt.read_particle_array()
    melt.close()
    acc, meas = {}, {}
    for prec in (0, 1):
        st = W.fcc_liquid(128, precision=prec)[1]
        sim = lj.LennardJones(state, st)
        m = sim.measure_now()
        ref = oracle_of(oracle, state, st)
        r = ref.measure()
        assert list(m.raw) == list(r.raw), f"precision {prec}: step-0 totals"
        _, _, a, o = sim.debug_state()
        _, _, ra, ro = ref.state()
        np.testing.assert_array_equal(o, ro)
        np.testing.assert_array_equal(a, ra)
        ref.close()
        acc[prec] = a
        meas[prec] = m
        sim.close()
    # the precision form against the f64 oracle (whose forces the FP64 form equals bit for bit)
    err = np.linalg.norm(acc[1] - acc[0], axis=1)
    rms = np.sqrt(np.mean(err ** 2) / np.mean(np.sum(acc[0] ** 2, axis=1)))
    pe = abs(meas[1].potential - meas[0].potential) / abs(meas[0].potential)
    w = abs(meas[1].virial - meas[0].virial) / abs(meas[0].virial)
    print(f"FCC-8M FP32 form: force rms relative {rms:.2e}, PE {pe:.2e}, W {w:.2e}")
    assert rms <= 1e-6
    assert pe <= 4e-7 and w <= 4e-7
    # NVE energy drift over 10^4 steps, both forms
    for prec in (0, 1):
        sim = lj.LennardJones(state, W.fcc_liquid(128, precision=prec)[1])
        e0 = sim.measure_now().energy()
        sim.run(10_000)
        m = sim.measure_now()
        drift = abs(m.energy() - e0) / abs(e0)
        print(f"FCC-8M precision {prec}: NVE drift over 10^4 steps {drift:.2e}")
        assert drift <= 1e-5
        assert m.list_overflow == 0
        sim.close()


def _chosen(meta, specs):
    """Smallest and largest N, densest and most dilute, hottest and coldest,
    and the smallest box (fewest cells: the workload has no replica below
    three cells per axis; two cells per axis are covered by
    test_gpu_api.py::test_two_cells_per_axis)."""
    n = np.array([m["n"] for m in meta])
    rho = np.array([m["rho"] for m in meta])
    temp = np.array([m["T"] for m in meta])
    side = np.array([st.box_size[0] for _, st in specs])
    picks = [int(n.argmin()), int(n.argmax()), int(rho.argmax()), int(rho.argmin()), int(temp.argmax()),
             int(temp.argmin()), int(side.argmin())]
    return sorted(set(picks))


def test_replica_batch_against_the_oracle(oracle, lj):
    specs, meta = W.batch_workload(512, 7)
    batch = lj.Batch(specs)
    ms = batch.measure_now()
    for r, (parts, st) in enumerate(specs):
        ref = oracle_of(oracle, parts, st)
        assert list(ms[r].raw) == list(ref.measure().raw), f"replica {r}: step-0 totals"
        ref.close()
    chosen = _chosen(meta, specs)
    assert len(chosen) >= 5
    batch.run(100)
    ms = batch.measure_now()
    allp = batch.read_particle_array()
    offs = np.concatenate([[0], np.cumsum([m["n"] for m in meta])])
    for r in chosen:
        parts, st = specs[r]
        ref = oracle_of(oracle, parts, st)
        ref.run(100)
        assert list(ms[r].raw) == list(ref.measure().raw), f"replica {r} ({meta[r]}): after 100 steps"
        pos, vel = ref.read_particles()
        np.testing.assert_array_equal(allp["pos"][offs[r]:offs[r + 1]], pos)
        np.testing.assert_array_equal(allp["vel"][offs[r]:offs[r + 1]], vel)
        ref.close()
    batch.run(4900)
    ms = batch.measure_now()
    assert max(m.list_overflow for m in ms) == 0
    cut = [batch.debug_replica_counters(r)["stage_overflow_builds"] for r in range(len(specs))]
    print(f"replicas with cut bricks: {sum(1 for c in cut if c)}; cut builds {sum(cut)}")
    # the replica whose bricks were cut most often, over the whole run
    r = int(np.argmax(cut))
    if cut[r] == 0:
        pytest.skip("no replica's bricks were cut in this run")
    parts, st = specs[r]
    ref = oracle_of(oracle, parts, st)
    ref.run(5000)
    assert list(ms[r].raw) == list(ref.measure().raw), f"replica {r} ({meta[r]}, {cut[r]} cut builds): 5000 steps"
    ref.close()


def test_slab_seed_bit_exact_with_a_profile_sample(oracle, lj):
    parts, box = W.slab_replica(77, 100)
    st = W.slab_settings(box, 1000)
    sim = lj.LennardJones(parts, st)
    ref = oracle_of(oracle, parts, st)
    assert list(sim.measure_now().raw) == list(ref.measure().raw)
    for steps in (1, 19):
        sim.run(steps)
        ref.run(steps)
        sim.sample_profile()
        ref.sample_profile()
        m, r = sim.measure_now(), ref.measure()
        assert list(m.raw) == list(r.raw), f"after {sim.steps()} steps"
    raw, samples = sim.read_profile_raw()
    r_raw, r_samples = ref.read_profile()
    assert samples == r_samples == 2
    np.testing.assert_array_equal(raw, r_raw)
    same_state(sim, ref)


def test_slab_batch_profiles_add_up(lj):
    specs = []
    for k in range(8):
        parts, box = W.slab_replica(77, 100 + 2 * k)
        specs.append((parts, W.slab_settings(box, 1000 + k)))
    batch = lj.Batch(specs)
    batch.run(20)
    batch.sample_profile()
    ms = batch.measure_now()
    for r, m in enumerate(ms):
        raw, samples = batch.read_profile_raw(r)
        assert samples == 1
        sums = raw[:-3].reshape(200, 8)
        assert int(sums[:, 0].sum()) == m.count << 32  # every atom in one slab
        ke = sums[:, 1:4].sum() / 2.0 ** 32 / 2.0
        assert abs(ke - m.kinetic) <= 1e-9 * abs(m.kinetic)
        # Irving-Kirkwood: each pair's virial spread over the slabs adds up to the pair's share
        w_slabs = sums[:, 4:7].sum(axis=0)
        w_box = raw[-3:]
        assert np.all(np.abs(w_slabs - w_box) <= 1e-9 * np.abs(w_box).max())
        # the profile's pair terms are FP64, the measurement's FP32 (the FP32 form)
        pe = sums[:, 7].sum() / 2.0 ** 32
        assert abs(pe - m.potential) <= 1e-6 * abs(m.potential)
        assert abs(w_box.sum() / 2.0 ** 32 - m.virial) <= 1e-6 * abs(m.virial)
        assert m.list_overflow == 0
