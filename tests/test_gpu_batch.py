"""Batched replicas on the GPU: replicas built like the batched benchmark
(FCC sites for any count, random insertion at low density) run in one
handle exactly as they do alone; an all-NVE batch (the drift-ahead step,
dense and sparse bricks on two streams) matches the oracle replica by
replica; and the dilute-gas check of the reference's
tests/thermodynamics.rs:234-293 holds for each replica of a batch, its
standard error coming from block means of the same run."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def potential(r2, rc2):
    """Truncated and shifted LJ (tests/thermodynamics.rs:20-31)."""
    inv6 = 1.0 / r2 ** 3
    c6 = 1.0 / rc2 ** 3
    return np.where(r2 < rc2, 4.0 * (inv6 * inv6 - inv6) - 4.0 * (c6 * c6 - c6), 0.0)


def second_virial(t, cutoff, steps=200_000):
    """B2 = -2 pi integral of (exp(-U/T) - 1) r^2 dr by Simpson's rule
    (tests/thermodynamics.rs:216-232)."""
    h = cutoff / steps
    r = np.arange(steps + 1) * h
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        f = np.where(r > 0, (np.exp(-potential(r * r, cutoff * cutoff) / t) - 1.0) * r * r, 0.0)
    w = np.where(np.arange(steps + 1) % 2 == 1, 4.0, 2.0)
    w[0] = w[-1] = 1.0
    return -2.0 * math.pi * float((f * w).sum()) * h / 3.0


def replica(lj, n, rho, temp, seed):
    if rho >= 0.3:
        parts, side = lj.fcc_count_array(n, rho, temp, seed)
    else:
        side = float(np.float32((n / rho) ** (1.0 / 3.0)))
        parts = lj.random_insertion_array(n, [side] * 3, 0.9, temp, seed)
    return parts, lj.Settings([side] * 3, 2.5, 0.005, skin=0.3, thermostat=lj.Langevin(temp, 1.0, seed))


def test_mixed_batch_equals_single_runs(lj):
    rng = np.random.default_rng(3)
    specs = []
    for _ in range(10):
        n = int(round(math.exp(rng.uniform(math.log(500), math.log(4000)))))
        specs.append(replica(lj, n, float(rng.uniform(0.01, 0.85)), float(rng.uniform(0.7, 2.0)),
                             int(rng.integers(1, 1 << 40))))
    batch = lj.Batch(specs)
    batch.run(80)
    ms = batch.measure_now()
    allp = batch.read_particle_array()
    off = 0
    for (parts, st), m in zip(specs, ms):
        single = lj.LennardJones(parts, st)
        single.run(80)
        s = single.measure_now()
        assert list(m.raw) == list(s.raw)
        n = single.count()
        exp = single.read_particle_array()
        np.testing.assert_array_equal(allp[off:off + n]["pos"], exp["pos"])
        np.testing.assert_array_equal(allp[off:off + n]["vel"], exp["vel"])
        off += n
        single.close()


def test_dilute_gas_virial_expansion_per_replica(lj):
    # tests/thermodynamics.rs:234-293 (N 4096, rho 0.01, T 2, Langevin gamma 1,
    # Z sampled every 20 steps, 20 block means), for four replicas at once
    cases = [(0.01, 2.0, 7), (0.01, 2.0, 8), (0.005, 1.6, 9), (0.005, 1.8, 10)]
    specs = []
    for rho, temp, seed in cases:
        side = (4096 / rho) ** (1.0 / 3.0)
        parts = lj.lattice_array(4096, side, temp, seed)
        specs.append((parts, lj.Settings([side] * 3, 2.5, 0.005, skin=0.3,
                                         thermostat=lj.Langevin(temp, 1.0, seed))))
    batch = lj.Batch(specs)
    batch.run(4000)
    z = []
    for _ in range(2000):
        batch.run(20)
        ms = batch.measure_now()
        z.append([m.pressure() / ((m.count / m.volume) * m.temperature(3.0 * m.count)) for m in ms])
    z = np.array(z)
    blocks = z.reshape(20, -1, len(cases)).mean(axis=1)
    mean = z.mean(axis=0)
    err = blocks.std(axis=0, ddof=1) / math.sqrt(20)
    for r, (rho, temp, _) in enumerate(cases):
        rho_box = 4096 / ms[r].volume
        predicted = 1.0 + second_virial(temp, 2.5) * rho_box
        print(f"replica {r}: rho {rho} T {temp}: Z = {mean[r]:.5f} +/- {err[r]:.5f}, 1 + B2 rho = {predicted:.5f}")
        assert abs(mean[r] - 1.0) < 0.01
        assert abs(mean[r] - predicted) < 4.0 * max(err[r], 1e-4)


def nve_replica(lj, n, rho, temp, seed, precision, shifted):
    parts, st = replica(lj, n, rho, temp, seed)
    return parts, lj.Settings(list(st.box_size), 2.5, 0.005, skin=0.3, precision=precision, shifted_force=shifted)


@pytest.mark.parametrize("precision,shifted", [(0, False), (1, True)])
def test_nve_batch_on_both_layouts_against_the_oracle(oracle, lj, precision, shifted):
    # Every replica NVE, so the batch steps in the drift-ahead form; the
    # dilute replicas' bricks take the sparse layout, whose launches run on
    # the side stream beside the dense ones. Each replica against the
    # oracle run alone, bit for bit, through list rebuilds: FP64 with the
    # plain cutoff, and the FP32 form with the shifted force.
    specs = [nve_replica(lj, n, rho, t, s, precision, shifted)
             for n, rho, t, s in [(2048, 0.8442, 0.9, 21), (1500, 0.05, 1.5, 22), (3000, 0.7, 1.2, 23),
                                  (900, 0.02, 2.0, 24)]]
    batch = lj.Batch(specs)
    c = batch.debug_counters()
    assert 0 < c["sparse_bricks"] < c["bricks"], c
    steps = 150
    for chunk in (1, 17, steps - 18):  # an eager first step, a graph of 16 and one of 1, then the rest
        batch.run(chunk)
    ms = batch.measure_now()
    allp = batch.read_particle_array()
    off = 0
    for r, (parts, st) in enumerate(specs):
        ref = oracle.Sim(np.ascontiguousarray(parts["pos"]), np.ascontiguousarray(parts["vel"]),
                         list(st.box_size), cutoff=2.5, skin=0.3, dt=0.005, shifted_force=shifted,
                         precision=precision)
        ref.run(steps)
        m = ref.measure()
        assert list(ms[r].raw) == list(m.raw), f"replica {r}"
        assert ms[r].list_rebuilds == m.list_rebuilds > 1, f"replica {r}"
        pos, vel = ref.read_particles()
        n = len(parts)
        np.testing.assert_array_equal(allp["pos"][off:off + n], pos)
        np.testing.assert_array_equal(allp["vel"][off:off + n], vel)
        off += n
        ref.close()
