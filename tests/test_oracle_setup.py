"""Initial states and derived parameters of the oracle.

Lattices follow lennard_jones.rs:1377-1441; the parameter rules follow
lennard_jones.rs:443-560 and are pinned by the sizes derived independently
from the reference's rules for the benchmark configurations, and by closed
forms."""
import math

import numpy as np
import pytest

from helpers import xorshift_stream


def test_lattice_sites_and_velocity_stream(oracle):
    n, box, T, seed = 1000, 11.0, 1.0, 3
    pos, vel = oracle.lattice(n, box, T, seed)
    gap = box / 10
    c = np.arange(n)
    expect = np.stack([(c % 10 + 0.5) * gap, ((c // 10) % 10 + 0.5) * gap, (c // 100 + 0.5) * gap], axis=1)
    np.testing.assert_array_equal(pos, expect)
    # velocities: xorshift64 (seed | 1), ((s >> 11) + 0.5) / 2^53, Box-Muller cosine branch,
    # then the mean is removed (lennard_jones.rs:1409-1439); checked on the first particle
    g = xorshift_stream(seed | 1)

    def r():
        return ((next(g) >> 11) + 0.5) / float(1 << 53)

    raw = []
    for _ in range(3 * n):
        a = math.sqrt(-2.0 * math.log(r()))
        raw.append(a * math.cos(math.tau * r()) * math.sqrt(T))
    raw = np.array(raw).reshape(n, 3)
    mean = np.array([sum(raw[:, k]) / n for k in range(3)])
    np.testing.assert_allclose(vel, raw - mean, rtol=0, atol=1e-15)
    assert np.all(np.abs(vel.sum(axis=0)) < 1e-11)
    ke = 0.5 * np.sum(vel ** 2)
    assert abs(ke / (1.5 * n * T) - 1) < 0.1


def test_lattice_box_fills_rectangular_box(oracle):
    size = [9.0, 10.5, 14.25]
    pos, vel = oracle.lattice_box(1000, size, 1.0, 3)
    for k in range(3):
        assert pos[:, k].min() > 0 and pos[:, k].max() < size[k]
    # distinct sites, same spacing rule on each axis
    assert len({tuple(p) for p in pos}) == 1000
    spacing = (np.prod(size) / 1000) ** (1 / 3)
    sites = [len(np.unique(pos[:, k])) for k in range(3)]
    for k in range(3):
        assert abs(size[k] / sites[k] - spacing) < 0.35 * spacing


def test_fcc_lattice(oracle):
    pos, vel, side = oracle.fcc(4, 0.8442, 0.722, 5)
    assert pos.shape == (256, 3)
    assert side == float(np.float32(side))
    assert abs(256 / side ** 3 - 0.8442) < 1e-5
    a = side / 4
    d = pos[:, None, :] - pos[None, :, :]
    d -= side * np.round(d / side)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(256) * 10
    assert abs(r.min() - a / math.sqrt(2)) < 1e-12
    assert np.all(np.sum(np.abs(r - a / math.sqrt(2)) < 1e-9, axis=1) == 12)
    assert np.all((pos >= 0) & (pos < side))
    assert np.all(np.abs(vel.sum(axis=0)) < 1e-11)


def test_random_insertion_keeps_min_distance(oracle):
    box = [12.0, 12.0, 12.0]
    pos, vel = oracle.random_insertion(300, box, 0.9, 1.2, 8)
    d = pos[:, None, :] - pos[None, :, :]
    d -= 12.0 * np.round(d / 12.0)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(300) * 10
    assert r.min() >= 0.9
    assert np.all(np.abs(vel.sum(axis=0)) < 1e-11)


def derive_for(oracle, n, box, **kw):
    pos = np.full((n, 3), 0.1) if n else np.zeros((0, 3))
    pos = np.random.default_rng(0).uniform(0, 1, (n, 3)) * np.array(box)
    vel = np.zeros((n, 3))
    return oracle.derive_params(pos, vel, box, **kw)


@pytest.mark.parametrize("n,box,L,cells,cap_ref", [
    (1000, [(1000 / 0.8442) ** (1 / 3)] * 3, [10.580793] * 3, [3, 3, 3], 188),
    (1048576, [(1048576 / 0.8442) ** (1 / 3)] * 3, [107.494156] * 3, [38, 38, 38], 188),
    (8388608, [(8388608 / 0.8442) ** (1 / 3)] * 3, [214.988312] * 3, [76, 76, 76], 188),
    (2000000, [136.2911, 136.2911, 545.1646], [136.2911, 136.2911, 545.1646], [48, 48, 194], 69),
])
def test_derived_grid_matches_the_configuration_table(oracle, n, box, L, cells, cap_ref):
    # sizes derived from lennard_jones.rs:450-474 for the benchmark configurations
    p = oracle.derive_params(np.zeros((1, 3)), np.zeros((1, 3)), box)
    for k in range(3):
        assert abs(p.box[k] - L[k]) < 1e-4
        assert p.box[k] == float(np.float32(box[k]))
    assert list(p.cells) == cells
    # capacity the reference would give at this count (box-average density)
    rho = n / (p.box[0] * p.box[1] * p.box[2])
    assert math.ceil(2 * 4 / 3 * math.pi * 2.8 ** 3 * rho + 32) == cap_ref


def test_cell_count_is_capped_at_2_pow_20(oracle):
    p = oracle.derive_params(np.zeros((1, 3)), np.zeros((1, 3)), [300.0] * 3)
    total = p.cells[0] * p.cells[1] * p.cells[2]
    assert total <= 1 << 20
    assert max(p.cells) - min(p.cells) <= 1
    assert all(math.prod([c + (k == a) for k, c in enumerate(p.cells)]) > 1 << 20 for a in range(3))


def test_pair_constants_closed_form(oracle):
    p = oracle.derive_params(np.zeros((1, 3)), np.zeros((1, 3)), [20.0] * 3, cutoff=2.5, shifted_force=True)
    rc = 2.5
    assert p.shift == pytest.approx(4 * (rc ** -12 - rc ** -6), rel=1e-15)
    assert p.force_shift == pytest.approx(24 * (2 * rc ** -12 - rc ** -6) / rc, rel=1e-15)
    assert p.rc2 == 6.25 and p.list_r2 == pytest.approx(2.8 ** 2, rel=1e-15)
    assert p.half_skin2 == pytest.approx(0.0225, rel=1e-15)
    q = oracle.derive_params(np.zeros((1, 3)), np.zeros((1, 3)), [20.0] * 3, thermostat=1, temperature=1.5,
                             friction=0.5, dt=0.004)
    assert q.c1 == pytest.approx(math.exp(-0.002), rel=1e-15)
    assert q.c2 == pytest.approx(math.sqrt((1 - math.exp(-0.004)) * 1.5), rel=1e-15)


def test_energy_scale_of_the_million_atom_liquid(oracle):
    pos, vel, side = oracle.fcc(64, 0.8442, 0.722, 2024)
    p = oracle.derive_params(pos, vel, [side] * 3)
    assert p.energy_scale == 0.5
    assert p.capacity >= p.capacity_reference


def test_box_narrower_than_two_list_radii_is_refused(oracle):
    with pytest.raises(ValueError, match="two"):
        oracle.derive_params(np.zeros((1, 3)), np.zeros((1, 3)), [5.5, 20.0, 20.0])
    p = oracle.derive_params(np.zeros((1, 3)), np.zeros((1, 3)), [5.61, 20.0, 20.0])
    assert p.cells[0] == 2
