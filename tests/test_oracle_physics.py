"""The oracle's cell grid, lists, forces and dynamics, pinned by brute force
over all pairs, closed forms and invariants (tests/thermodynamics.rs:20-98,
tests/coexistence.rs:38-130)."""
import math

import numpy as np
import pytest

from helpers import jittered_box, jittered_lattice, sc_liquid


def wrapped(pos, box):
    return np.stack([np.mod(pos[:, k], box[k]) for k in range(3)], axis=1)


def test_cell_grid_finds_every_pair(oracle):
    # tests/thermodynamics.rs:55-98, at 1e-12 instead of the reference's 1e-5
    pos, vel = jittered_lattice(oracle)
    s = oracle.Sim(pos, vel, [11.0] * 3, dt=0.002)
    m = s.measure()
    L = list(s.params.box)
    pe, w = oracle.all_pairs(wrapped(pos, L), L, 2.5, False)
    assert abs((m.potential - pe) / pe) < 1e-12
    assert abs((m.virial - w.sum()) / w.sum()) < 1e-12


@pytest.mark.parametrize("shifted_force", [False, True])
def test_rectangular_box_matches_all_pairs(oracle, shifted_force):
    # tests/coexistence.rs:93-130
    size = [9.0, 10.5, 14.25]
    pos, vel = jittered_box(oracle, 1000, size)
    s = oracle.Sim(pos, vel, size, shifted_force=shifted_force)
    m = s.measure()
    box = list(s.params.box)
    pe, w, f = oracle.all_pairs(wrapped(pos, box), box, 2.5, shifted_force, forces=True)
    assert abs((m.potential - pe) / pe) < 1e-12
    assert abs((m.virial - w.sum()) / w.sum()) < 1e-12
    _, _, acc, orig = s.state()
    F = np.zeros_like(acc)
    F[orig] = acc
    assert np.max(np.abs(F - f)) < 1e-10 * np.max(np.abs(f))
    assert np.all(np.abs(acc.sum(axis=0)) < 1e-10)


def test_two_body_closed_form(oracle):
    for sf in (False, True):
        for r in (0.95, 1.12, 1.7, 2.45):
            pos = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
            s = oracle.Sim(pos, np.zeros((2, 3)), [12.0] * 3, shifted_force=sf)
            m = s.measure()
            rc = 2.5
            u = 4 * (r ** -12 - r ** -6) - 4 * (rc ** -12 - rc ** -6)
            rf = 24 * (2 * r ** -12 - r ** -6)
            if sf:
                fc = 24 * (2 * rc ** -12 - rc ** -6) / rc
                u += (r - rc) * fc
                rf -= r * fc
            assert m.potential == pytest.approx(u, rel=1e-12, abs=1e-12)
            assert m.virial == pytest.approx(rf, rel=1e-12, abs=1e-12)
            _, _, acc, orig = s.state()
            F = np.zeros_like(acc)
            F[orig] = acc
            assert F[1, 0] == pytest.approx(rf / r, rel=1e-12)
            assert F[0, 0] == pytest.approx(-rf / r, rel=1e-12)


def test_cell_order_and_lists(oracle):
    pos, vel, box = sc_liquid(oracle, n_side=9)
    pos = oracle.jitter(pos, 99, 0.4)
    s = oracle.Sim(pos, vel, [box] * 3)
    s.prime()
    p = s.params
    x, _, _, _ = s.state()
    cell_of, start, sorted_ = s.cells()
    n = x.shape[0]
    # R2 restated in numpy f32: every atom in the cell its position gives
    cs = np.array(p.cell_size, dtype=np.float32)
    c3 = np.clip(np.floor(x.astype(np.float32) / cs).astype(np.int64), 0, np.array(p.cells) - 1)
    np.testing.assert_array_equal(cell_of, (c3[:, 2] * p.cells[1] + c3[:, 1]) * p.cells[0] + c3[:, 0])
    # every rebuild leaves the atoms in cell order: sorted is the identity
    np.testing.assert_array_equal(sorted_, np.arange(n))
    assert start[-1] == n and np.all(np.diff(start) >= 0)
    assert np.all(np.diff(cell_of.astype(np.int64)) >= 0)
    # list membership equals brute force within rc + skin (nearest image)
    counts, entries = s.lists()
    L = np.array(p.box)
    d = x[None, :, :] - x[:, None, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    for i in range(n):
        got = entries[i, :counts[i]]
        assert set(got.tolist()) == set(np.nonzero(r2[i] < p.list_r2)[0].tolist())
        # order: stencil cell (dz, dy, dx) of each entry, then slot, ascending
        ci = np.array([cell_of[i] % p.cells[0], (cell_of[i] // p.cells[0]) % p.cells[1],
                       cell_of[i] // (p.cells[0] * p.cells[1])])
        cj = np.stack([cell_of[got] % p.cells[0], (cell_of[got] // p.cells[0]) % p.cells[1],
                       cell_of[got] // (p.cells[0] * p.cells[1])], axis=1)
        off = (cj - ci + 1) % np.array(p.cells)  # 0, 1, 2 for -1, 0, +1
        key = ((off[:, 2] * 3 + off[:, 1]) * 3 + off[:, 0]) * n + got
        assert np.all(np.diff(key) > 0)


def test_fp32_pair_forces_error_against_fp64(oracle):
    pos, vel, box = sc_liquid(oracle, n_side=12)
    pos = oracle.jitter(pos, 99, 0.3)
    a = oracle.Sim(pos, vel, [box] * 3, precision=0)
    b = oracle.Sim(pos, vel, [box] * 3, precision=1)
    ma, mb = a.measure(), b.measure()
    fa, fb = a.state()[2], b.state()[2]
    rel_rms = math.sqrt(np.mean(np.sum((fb - fa) ** 2, axis=1)) / np.mean(np.sum(fa ** 2, axis=1)))
    assert rel_rms < 1e-6
    assert abs((mb.potential - ma.potential) / ma.potential) < 4e-7
    assert abs((mb.virial - ma.virial) / ma.virial) < 4e-7


def test_nve_energy_error_is_second_order_and_momentum_is_conserved(oracle):
    # velocity Verlet: the energy error over a fixed time falls as dt^2
    pos, vel, box = sc_liquid(oracle)
    errs = []
    for dt in (0.005, 0.0025):
        s = oracle.Sim(pos, vel, [box] * 3, shifted_force=True, dt=dt)
        e0 = s.measure()
        s.run(int(round(1.0 / dt)))
        e1 = s.measure()
        errs.append(abs((e1.kinetic + e1.potential - e0.kinetic - e0.potential) / (e0.kinetic + e0.potential)))
        assert max(abs(p) for p in e1.momentum) < 1e-9
        assert e1.list_overflow == 0 and e1.list_rebuilds > 1
    assert errs[1] < 1e-4
    assert 2.5 < errs[0] / errs[1] < 6.0


def test_lists_always_hold_every_pair_within_the_cutoff(oracle):
    # the skin rule (lennard_jones.wgsl:20-25): between rebuilds no pair within rc is missing
    pos, vel, box = sc_liquid(oracle, n_side=8, temperature=2.0)
    s = oracle.Sim(pos, vel, [box] * 3)
    L = np.array(s.params.box)
    for _ in range(12):
        s.run(7)
        x = s.state()[0]
        counts, entries = s.lists()
        d = x[None, :, :] - x[:, None, :]
        d -= L * np.round(d / L)
        r2 = (d ** 2).sum(-1)
        np.fill_diagonal(r2, np.inf)
        for i in range(x.shape[0]):
            assert set(np.nonzero(r2[i] < 6.25)[0].tolist()) <= set(entries[i, :counts[i]].tolist())


def test_threads_and_reruns_give_identical_bits(oracle):
    pos, vel, box = sc_liquid(oracle, n_side=8)
    out = []
    for threads in (1, oracle.max_threads(), 1):
        oracle.set_threads(threads)
        s = oracle.Sim(pos, vel, [box] * 3, thermostat=1, temperature=1.5, friction=0.5, seed=3)
        s.run(150)
        m = s.measure()
        out.append((s.state()[0].tobytes(), s.state()[1].tobytes(), list(m.raw)))
    oracle.set_threads(oracle.max_threads())
    assert out[0] == out[1] == out[2]


def test_time_reversibility(oracle):
    pos, vel, box = sc_liquid(oracle, n_side=8, temperature=0.9)
    s = oracle.Sim(pos, vel, [box] * 3, shifted_force=True, dt=0.002)
    s.run(100)
    p1, v1 = s.read_particles()
    back = oracle.Sim(p1, -v1, [box] * 3, shifted_force=True, dt=0.002)
    back.run(100)
    p2, v2 = back.read_particles()
    L = s.params.box[0]
    d = p2 - np.mod(pos, L)
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-8
    assert np.max(np.abs(v2 + vel)) < 1e-8


def test_langevin_holds_the_temperature(oracle):
    n, T = 512, 1.5
    box = (n / 0.3) ** (1 / 3)
    pos, vel = oracle.lattice(n, box, 0.5, 21)
    s = oracle.Sim(pos, vel, [box] * 3, thermostat=1, temperature=T, friction=1.0, seed=9)
    s.run(1000)
    temps = []
    for _ in range(40):
        s.run(25)
        temps.append(2 * s.measure().kinetic / (3 * n))
    mean = np.mean(temps)
    assert abs(mean / T - 1) < 0.03


def nh_sim(oracle, n=512, T=1.5, start_T=0.5, period=0.5, seed=21):
    box = (n / 0.3) ** (1 / 3)
    pos, vel = oracle.lattice(n, box, start_T, seed)
    return oracle.Sim(pos, vel, [box] * 3, dt=0.002, shifted_force=True, thermostat=2, temperature=T,
                      period=period), n


def test_nose_hoover_parameters(oracle):
    # lennard_jones.rs:516-528: g = 3N - 3, Q = g k T period^2 / (4 pi^2);
    # set_temperature keeps the period (rs:912-920)
    s, n = nh_sim(oracle)
    g = 3 * n - 3
    assert s.params.nh_gkt == g * 1.5
    assert s.params.nh_q == pytest.approx(g * 1.5 * 0.25 / (4 * np.pi ** 2), rel=1e-15)
    s.set_temperature(0.75)
    assert s.params.nh_gkt == g * 0.75
    assert s.params.nh_q == pytest.approx(g * 0.75 * 0.25 / (4 * np.pi ** 2), rel=1e-15)


def test_nose_hoover_conserves_the_extended_energy(oracle):
    # KE + PE + Q xi^2/2 + g k T eta is the conserved quantity of the
    # splitting: its largest deviation over the same 4 time units falls as
    # dt^2 (a wrong sign or a missing term would leave an O(1) drift), and
    # the thermostat heats a cold start to the target on average
    n = 512
    box = (n / 0.3) ** (1 / 3)
    pos, vel = oracle.lattice(n, box, 0.5, 21)
    s = oracle.Sim(pos, vel, [box] * 3, dt=0.002, shifted_force=True)
    s.run(1000)
    p, v = s.read_particles()
    devs, temps = [], []
    for dt in (0.004, 0.002):
        t = oracle.Sim(p, v, [box] * 3, dt=dt, shifted_force=True, thermostat=2, temperature=1.5, period=0.5)
        m0 = t.measure()
        e0 = m0.kinetic + m0.potential + m0.thermostat_energy
        dev = 0.0
        every = int(round(0.2 / dt))
        for _ in range(20):
            t.run(every)
            m = t.measure()
            dev = max(dev, abs(m.kinetic + m.potential + m.thermostat_energy - e0))
            temps.append(2 * m.kinetic / (3 * n - 3))
        devs.append(dev)
    assert 3.0 < devs[0] / devs[1] < 5.0
    assert abs(np.mean(temps) / 1.5 - 1) < 0.15


def test_nose_hoover_half_step_closed_form(oracle):
    # with no pair inside the cutoff the velocities only feel the thermostat:
    # two particles far apart, KE fixed by hand; one step scales v by
    # s1 = exp(-xi1 dt/2) twice, xi following the rule of R13
    box = [12.0] * 3
    pos = np.array([[1.0, 1.0, 1.0], [7.0, 7.0, 7.0]])
    vel = np.array([[0.3, -0.2, 0.1], [-0.3, 0.2, -0.1]])
    T, period, dt = 1.0, 0.5, 0.002
    s = oracle.Sim(pos, vel, box, dt=dt, thermostat=2, temperature=T, period=period)
    g = 3 * 2 - 3
    gkt, Q = g * T, g * T * period ** 2 / (4 * np.pi ** 2)
    ke = 0.5 * np.sum(vel ** 2)
    xi = eta = 0.0

    def half(ke, xi, eta):
        xi += (2 * ke - gkt) / Q * dt / 4
        s = np.exp(-xi * dt / 2)
        eta += xi * dt / 2
        xi += (2 * ke * s * s - gkt) / Q * dt / 4
        return s, xi, eta

    s1, xi, eta = half(ke, xi, eta)
    ke *= s1 * s1
    s2, xi, eta = half(ke, xi, eta)
    s.run(1)
    _, v = s.read_particles()
    np.testing.assert_allclose(v, vel * s1 * s2, rtol=1e-12)
    m = s.measure()
    assert m.thermostat_energy == pytest.approx(0.5 * Q * xi * xi + gkt * eta, rel=1e-9)
