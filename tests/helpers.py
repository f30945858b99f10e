"""Shared inputs of the tests, built the way the reference's tests build them."""
import numpy as np


def xorshift_stream(seed):
    """The reference tests' xorshift64 generator (tests/thermodynamics.rs:61-67)."""
    s = seed
    mask = (1 << 64) - 1
    while True:
        s ^= (s << 13) & mask
        s ^= s >> 7
        s ^= (s << 17) & mask
        yield s


def jittered_lattice(oracle, count=1000, box=11.0, temperature=1.0, seed=3, jitter_seed=99):
    """lattice(count, box, T, seed) jittered by +-0.15 (tests/thermodynamics.rs:60-69)."""
    pos, vel = oracle.lattice(count, box, temperature, seed)
    return oracle.jitter(pos, jitter_seed, 0.3), vel


def jittered_box(oracle, count, size, temperature=1.0, seed=3, jitter_seed=99):
    """lattice_box jittered by +-0.15 (tests/coexistence.rs:78-91)."""
    pos, vel = oracle.lattice_box(count, size, temperature, seed)
    return oracle.jitter(pos, jitter_seed, 0.3), vel


def sc_liquid(oracle, n_side=10, density=0.8442, temperature=0.722, seed=1):
    """Simple cubic n^3 at density rho, Maxwell-Boltzmann velocities, zero momentum."""
    n = n_side ** 3
    box = (n / density) ** (1.0 / 3.0)
    pos, vel = oracle.lattice(n, box, temperature, seed)
    return pos, vel, box


def spread_values(n):
    """tests/gpu.rs:75-87: 2^20 values from 1e-6 to 1e3 (a few to 1e6), both
    signs, as f32; random(7) is xorshift64 seeded with 7 (tests/gpu.rs:18-27)."""
    g = xorshift_stream(7)

    def r():
        return (next(g) >> 11) / float(1 << 53)

    out = np.zeros(n, dtype=np.float32)
    for i in range(n):
        top = 6.0 if i % 1024 == 0 else 3.0
        magnitude = 10.0 ** (r() * (top + 6.0) - 6.0)
        out[i] = np.float32(-magnitude if r() < 0.5 else magnitude)
    return out
