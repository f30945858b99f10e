"""The synthetic workloads of the benchmark and of the configuration tests.

Each builder returns host particles (the structured PARTICLE_DTYPE array)
and Settings, built with the library's host lattices, so the benchmark, the
tests and a user reproducing a number start from the same bits.

  sc_liquid      N = n^3 on a simple cubic lattice (lattice(), lennard_jones.rs:1379)
  fcc_liquid     FCC 4 n^3 at rho 0.8442, T 0.722 (4 x 64^3 = 1,048,576; 4 x 128^3 = 8,388,608)
  batch_workload replicas of log-uniform N in [500, 50000], rho in [0.01, 0.85],
                 T in [0.7, 2.0], Langevin gamma 1
  slab_replica   a liquid-vapour slab, Lz = 4 Lx, FCC liquid in the middle quarter
"""
from __future__ import annotations

import math

import numpy as np

from . import lennard_jones as lj

DENSITY = 0.8442
TEMPERATURE = 0.722
FCC_SEED = 2024


def sc_liquid(n_side: int = 10, density: float = DENSITY, temperature: float = TEMPERATURE, seed: int = 1):
    """n^3 particles on a simple cubic lattice at `density` with
    Maxwell-Boltzmann velocities (zero momentum); rc 2.5 truncated and
    shifted, skin 0.3, NVE, dt 0.005, FP64. Returns (particles, settings)."""
    n = n_side ** 3
    box = (n / density) ** (1.0 / 3.0)
    parts = lj.lattice_array(n, box, temperature, seed)
    return parts, lj.Settings([box] * 3, 2.5, 0.005, skin=0.3)


def fcc_liquid(cells: int, precision: int = lj.FP64, seed: int = FCC_SEED):
    """FCC 4 cells^3 at rho 0.8442, T 0.722 (sites offset a/4); rc 2.5
    truncated and shifted, skin 0.3, NVE, dt 0.005. Returns (particles,
    settings)."""
    parts, side = lj.fcc_array(cells, DENSITY, TEMPERATURE, seed)
    return parts, lj.Settings([side] * 3, 2.5, 0.005, skin=0.3, precision=precision)


def batch_workload(replicas: int = 512, seed: int = 7):
    """The batched-replica workload: N = round(exp(U(ln 500, ln 50000))),
    rho ~ U(0.01, 0.85), T ~ U(0.7, 2.0); FCC sites for rho >= 0.3, random
    insertion 0.9 apart below; Langevin gamma 1 with a seed per replica;
    rc 2.5, skin 0.3, dt 0.005, FP64. Returns ([(particles, settings)],
    [{"n", "rho", "T", "seed"}])."""
    rng = np.random.default_rng(seed)
    specs, meta = [], []
    for _ in range(replicas):
        n = int(round(math.exp(rng.uniform(math.log(500.0), math.log(50000.0)))))
        rho = float(rng.uniform(0.01, 0.85))
        temp = float(rng.uniform(0.7, 2.0))
        rseed = int(rng.integers(1, 1 << 62))
        if rho >= 0.3:
            parts, side = lj.fcc_count_array(n, rho, temp, rseed)
        else:
            side = float(np.float32((n / rho) ** (1.0 / 3.0)))
            parts = lj.random_insertion_array(n, [side] * 3, 0.9, temp, rseed)
        st = lj.Settings([side] * 3, 2.5, 0.005, skin=0.3, thermostat=lj.Langevin(temp, 1.0, rseed))
        specs.append((parts, st))
        meta.append({"n": n, "rho": n / side ** 3, "T": temp, "seed": rseed})
    return specs, meta


def slab_replica(cells: int, seed: int):
    """The liquid-vapour slab: an FCC liquid block of cells^3 unit cells at
    rho ~0.73 spanning the middle quarter of a box with Lz = 4 Lx, random
    vapour at rho 0.02 (0.9 apart, 1 sigma clear of the liquid) elsewhere;
    T 0.7 velocities, zero total momentum. Returns (particles, box)."""
    liquid, lx = lj.fcc_array(cells, 0.73, 0.7, seed)
    lz = 4.0 * lx
    liquid["pos"][:, 2] += 1.5 * lx  # the block fills z in [1.5 Lx, 2.5 Lx)
    gap = 1.0
    height = 3.0 * lx - 2.0 * gap
    n_vap = int(round(0.02 * lx * lx * height))
    vapour = lj.random_insertion_array(n_vap, [lx, lx, height], 0.9, 0.7, seed + 1)
    vapour["pos"][:, 2] = np.mod(vapour["pos"][:, 2] + 2.5 * lx + gap, lz)
    parts = np.concatenate([liquid, vapour])
    parts["vel"] = parts["vel"] - parts["vel"].mean(axis=0)
    return parts, [lx, lx, lz]


def slab_settings(box, seed: int, bins: int = 200):
    """The slab's settings: force-shifted rc 2.5, skin 0.3, dt 0.005,
    Langevin gamma 0.5 at T 0.7, FP32 pair forces, `bins` z slabs."""
    return lj.Settings(box, 2.5, 0.005, skin=0.3, shifted_force=True, profile_bins=bins, precision=lj.FP32,
                       thermostat=lj.Langevin(0.7, 0.5, seed))
