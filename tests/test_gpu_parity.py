"""GPU parity: the HIP library against the oracle on the same inputs.

Integer work (cells, cell order, lists, fixed-point sums) must be bit-exact.
FP64 forces, energies and trajectories follow the same IEEE operation
sequence as the oracle, so they are compared bit for bit as well. The FP32
pair-force form follows the oracle's R5f rule bit for bit too, and is
compared with the FP64 oracle within stated tolerances.
"""
import numpy as np
import pytest

from helpers import jittered_box, jittered_lattice, sc_liquid, spread_values

pytestmark = pytest.mark.gpu


def make_pair(oracle, lj, pos, vel, box, cutoff=2.5, skin=0.3, dt=0.005, shifted_force=False, precision=0,
              thermostat=None):
    box3 = [box] * 3 if np.isscalar(box) else list(box)
    settings = lj.Settings(box3, cutoff, dt, thermostat=thermostat, shifted_force=shifted_force, skin=skin,
                           precision=precision)
    sim = lj.LennardJones((pos, vel), settings)
    kind, temp, fric, seed = 0, 0.0, 0.0, 1
    if thermostat is not None:
        kind, temp, fric, seed = 1, thermostat.temperature, thermostat.friction, thermostat.seed
    ref = oracle.Sim(pos, vel, box3, cutoff=cutoff, skin=skin, dt=dt, shifted_force=shifted_force,
                     precision=precision, thermostat=kind, temperature=temp, friction=fric, seed=seed)
    return sim, ref


def assert_same_grid(sim, ref):
    g = sim.debug_grid()
    p = ref.params
    assert g["cells"] == list(p.cells)
    assert g["capacity"] == p.capacity
    cell_of, start, sorted_ = sim.debug_cells()
    r_cell_of, r_start, r_sorted = ref.cells()
    np.testing.assert_array_equal(cell_of, r_cell_of)
    np.testing.assert_array_equal(start, r_start)
    np.testing.assert_array_equal(sorted_, r_sorted)
    counts, entries = sim.debug_lists()
    r_counts, r_entries = ref.lists()
    np.testing.assert_array_equal(counts, r_counts)
    for i in range(len(counts)):
        np.testing.assert_array_equal(entries[i, : counts[i]], r_entries[i, : r_counts[i]])


def test_fixed_point_sum_kat(lj):
    # tests/gpu.rs:89-149: 2^20 spread values, exact integer total, same bits every run
    values = spread_values(1 << 20).astype(np.float64)
    expected = sum(int(round(float(x) * 2.0 ** 32)) for x in values)
    runs = [lj.fixed_sum(values, 2.0 ** 32) for _ in range(3)]
    assert runs == [expected] * 3


def test_force_reciprocal_rounds_like_division(lj):
    # the force loop's 1 / r^2 must be the correctly rounded quotient that
    # the oracle's 1.0 / r2 gives, over the whole range r^2 takes
    rng = np.random.default_rng(5)
    x = np.concatenate([
        10.0 ** rng.uniform(-4, 37, 4_000_000),
        rng.uniform(0.5, 8.0, 2_000_000),                       # liquid pair distances
        2.0 ** np.arange(-13.0, 122.0),                          # exact powers of two
        np.nextafter(2.0 ** np.arange(-13.0, 122.0), np.inf),
        np.nextafter(2.0 ** np.arange(-13.0, 122.0), 0.0),
        1.0 + np.arange(1, 20000) * 2.0 ** -52,                  # just above 1
        1.0 - np.arange(1, 20000) * 2.0 ** -53,                  # just below 1
        np.array([3e36, 6.25, 12.25, 7.84, 1e-4]),
    ])
    got = lj.debug_rcp(x)
    want = 1.0 / x
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} differ, first x = {x[bad[:5]]!r}"


def test_fp32_reciprocal_and_root_round_like_ieee(lj):
    # R5f's 1.0f / q and sqrtf(q): every float of two binades (the rules
    # scale by powers of two) and the range q = (float)r^2 takes
    rng = np.random.default_rng(11)
    binades = [((np.uint32(e) << np.uint32(23)) | np.arange(1 << 23, dtype=np.uint32)).view(np.float32)
               for e in (126, 127)]
    x = np.concatenate(binades + [
        (10.0 ** rng.uniform(-4, 36, 2_000_000)).astype(np.float32),
        np.float32(2.0) ** np.arange(-13, 120, dtype=np.float32),
        np.array([3e36, 6.25, 12.25, 7.84, 1e-4], dtype=np.float32),
    ])
    rcp, root = lj.debug_rcp32(x)
    want = np.float32(1.0) / x
    bad = np.flatnonzero(rcp.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"reciprocal: {bad.size} differ, first x = {x[bad[:5]]!r}"
    want = np.sqrt(x)
    bad = np.flatnonzero(root.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"square root: {bad.size} differ, first x = {x[bad[:5]]!r}"


@pytest.mark.parametrize("shifted_force", [False, True])
def test_fp32_form_bit_exact_with_the_oracle(oracle, lj, shifted_force):
    # R5f on both sides: the FP32 pair forces, summed in FP32 per list
    # octet, and the FP32 pair energies give the oracle's bits
    pos, vel, box = sc_liquid(oracle, n_side=12)
    pos = oracle.jitter(pos, 99, 0.3)
    sim, ref = make_pair(oracle, lj, pos, vel, box, precision=1, shifted_force=shifted_force)
    for steps in (0, 7, 60):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert list(m.raw) == list(r.raw), f"after {sim.steps()} steps"
    p, v, a, o = sim.debug_state()
    rp, rv, ra, ro = ref.state()
    np.testing.assert_array_equal(o, ro)
    np.testing.assert_array_equal(a, ra)
    np.testing.assert_array_equal(p, rp)
    np.testing.assert_array_equal(v, rv)


def test_cell_grid_finds_every_pair(oracle, lj):
    # tests/thermodynamics.rs:55-98, at the reference's tolerance and bit for bit
    pos, vel = jittered_lattice(oracle)
    sim, ref = make_pair(oracle, lj, pos, vel, 11.0, dt=0.002)
    m = sim.measure_now()
    r = ref.measure()
    assert_same_grid(sim, ref)
    assert list(m.raw) == list(r.raw)
    L = sim.box_size()[0]
    pe, w = oracle.all_pairs(np.mod(pos, L), [L] * 3, 2.5, False)
    assert abs((m.potential - pe) / pe) < 4e-7
    assert abs((m.virial - w.sum()) / w.sum()) < 4e-7
    _, _, acc, orig = sim.debug_state()
    _, _, r_acc, r_orig = ref.state()
    np.testing.assert_array_equal(orig, r_orig)
    np.testing.assert_array_equal(acc, r_acc)


@pytest.mark.parametrize("shifted_force", [False, True])
def test_rectangular_box_matches_all_pairs(oracle, lj, shifted_force):
    # tests/coexistence.rs:93-130 (scalar virial; the tensor comes with profiles)
    size = [9.0, 10.5, 14.25]
    pos, vel = jittered_box(oracle, 1000, size)
    sim, ref = make_pair(oracle, lj, pos, vel, size, shifted_force=shifted_force)
    m = sim.measure_now()
    r = ref.measure()
    assert_same_grid(sim, ref)
    assert list(m.raw) == list(r.raw)
    box = sim.box_size()
    wrapped = np.stack([np.mod(pos[:, k], box[k]) for k in range(3)], axis=1)
    pe, w = oracle.all_pairs(wrapped, box, 2.5, shifted_force)
    assert abs((m.potential - pe) / pe) < 4e-7
    assert abs((m.virial - w.sum()) / w.sum()) < 4e-7


def test_nve_trajectory_bit_exact(oracle, lj):
    # SC 10^3 at rho 0.8442, T 0.722, rc 2.5 truncated-shifted, dt 0.005:
    # the SC-1000 NVE configuration over its full 1000 steps
    pos, vel, box = sc_liquid(oracle)
    sim, ref = make_pair(oracle, lj, pos, vel, box)
    for steps in (1, 9, 90, 900):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert list(m.raw) == list(r.raw), f"after {sim.steps()} steps"
        assert m.list_rebuilds == r.list_rebuilds
    p, v, a, o = sim.debug_state()
    rp, rv, ra, ro = ref.state()
    np.testing.assert_array_equal(o, ro)
    np.testing.assert_array_equal(p, rp)
    np.testing.assert_array_equal(v, rv)
    np.testing.assert_array_equal(a, ra)
    assert_same_grid(sim, ref)


def test_fp32_forces_against_fp64_oracle(oracle, lj):
    pos, vel, box = sc_liquid(oracle, n_side=12)
    pos = oracle.jitter(pos, 99, 0.3)
    sim, _ = make_pair(oracle, lj, pos, vel, box, precision=1)
    ref = oracle.Sim(pos, vel, [box] * 3, precision=0)
    m, r = sim.measure_now(), ref.measure()
    _, _, acc, orig = sim.debug_state()
    _, _, r_acc, r_orig = ref.state()
    np.testing.assert_array_equal(orig, r_orig)
    err = np.linalg.norm(acc - r_acc, axis=1)
    rel_rms = np.sqrt(np.mean(err ** 2) / np.mean(np.sum(r_acc ** 2, axis=1)))
    print(f"FP32 forces: rms relative error {rel_rms:.2e}; PE {(m.potential - r.potential) / r.potential:.2e}; "
          f"W {(m.virial - r.virial) / r.virial:.2e}")
    assert rel_rms < 1e-6
    assert abs((m.potential - r.potential) / r.potential) < 4e-7
    assert abs((m.virial - r.virial) / r.virial) < 4e-7


@pytest.mark.parametrize("rho", [0.01, 0.05, 0.2])
def test_dilute_lists_match_the_oracle(oracle, lj, rho):
    # dilute systems (sparse bricks, cells of one or a few atoms, items of
    # one atom): every list holds the oracle's entries in the oracle's
    # order, before and after rebuilds
    n = 3000
    side = float(np.float32((n / rho) ** (1.0 / 3.0)))
    parts = lj.random_insertion_array(n, [side] * 3, 0.9, 1.5, 29)
    pos, vel = np.ascontiguousarray(parts["pos"]), np.ascontiguousarray(parts["vel"])
    sim, ref = make_pair(oracle, lj, pos, vel, side)
    assert list(sim.measure_now().raw) == list(ref.measure().raw)  # first lists built
    assert_same_grid(sim, ref)
    sim.run(200)
    ref.run(200)
    assert sim.debug_counters()["rebuilds"] > 1
    assert_same_grid(sim, ref)
    assert list(sim.measure_now().raw) == list(ref.measure().raw)


def test_neighbourhood_past_8192_atoms_keeps_exact_lists(oracle, lj):
    # cutoff 7 at rho 0.85: 3 cells per axis, so every one-cell brick's 27
    # neighbour cells hold all 9261 atoms, beyond LDS and beyond what a
    # list entry of 8 x the staged index can hold in 16 bits; the global
    # path stores the staged index itself and the lists stay exact
    n = 21 ** 3
    box = (n / 0.85) ** (1.0 / 3.0)
    pos, vel = oracle.lattice(n, box, 0.722, 4)
    pos = oracle.jitter(pos, 7, 0.2)
    sim, ref = make_pair(oracle, lj, pos, vel, box, cutoff=7.0, dt=0.002)
    assert sim.debug_grid()["cells"] == [3, 3, 3]
    for steps in (0, 20):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert m.list_overflow == r.list_overflow == 0
        assert list(m.raw) == list(r.raw)
    assert sim.debug_counters()["global_cell_builds"] > 0
    assert_same_grid(sim, ref)
    np.testing.assert_array_equal(sim.debug_state()[2], ref.state()[2])


def test_list_overflow_matches_the_oracle(oracle, lj):
    # lennard_jones.wgsl:494-497: past the capacity a list keeps its first
    # `capacity` entries and the measurement reports the longest list found;
    # a capacity of 40 (the rule gives ~190 here) overflows every list
    pos, vel = jittered_lattice(oracle)
    with lj.debug_plan(list_capacity=40), oracle.capacity_override(40):
        sim, ref = make_pair(oracle, lj, pos, vel, 11.0, dt=0.002)
    for steps in (0, 50):
        sim.run(steps)
        ref.run(steps)
        m, r = sim.measure_now(), ref.measure()
        assert r.list_overflow > 40
        assert m.list_overflow == r.list_overflow
        assert list(m.raw) == list(r.raw)
        assert_same_grid(sim, ref)
