"""The oracle's fixed point against the reference's own known-answer tests
(crates/virial-precision/src/fixed.rs:69-113, tests/gpu.rs:89-149) and the
closed form round-ties-even(x * 2^32)."""
import math

import numpy as np
import pytest

from helpers import spread_values, xorshift_stream

ONE = 2.0 ** 32


def test_round_trips_values_with_no_bits_below_the_resolution(oracle):
    # fixed.rs:73-80
    for x in [0.0, 1.0, -1.0, 0.5, -0.5, 1234.5678, -98765.43, 0.0625, -3.0e5]:
        xf = float(np.float32(x))
        assert oracle.fixed_to_raw(xf) / ONE == xf


def test_rounds_finer_bits_to_the_nearest_step(oracle):
    # fixed.rs:82-88
    for x in [1e-3, -1e-3, 1e-12, -7.25e-9, 3.3e-6]:
        xf = float(np.float32(x))
        assert abs(oracle.fixed_to_raw(xf) / ONE - xf) <= 0.5 / ONE


def test_matches_exact_rounding_across_magnitudes(oracle):
    # fixed.rs:90-103: 200,000 xorshift magnitudes from 2^-34 to 2^30
    g = xorshift_stream(12345)
    for _ in range(200_000):
        seed = next(g)
        magnitude = 2.0 ** ((seed >> 11) / float(1 << 53) * 64.0 - 34.0)
        x = float(np.float32(magnitude if seed & 1 == 0 else -magnitude))
        exact = round(x * ONE)  # exact product in f64, Python rounds ties to even
        assert oracle.fixed_to_raw(x) == exact, x


def test_split_words_are_twos_complement(oracle):
    # fixed.rs:50-52: (low, high) join into the signed 64-bit total
    for x in [-1.0, -0.5, -1e-9, 3.75, -2.0 ** 30]:
        lo, hi = oracle.fixed_split(float(np.float32(x)))
        raw = ((hi << 32) | lo)
        if raw >= 1 << 63:
            raw -= 1 << 64
        assert raw == oracle.fixed_to_raw(float(np.float32(x)))


def test_sums_are_exact_in_any_order(oracle):
    # fixed.rs:105-113
    values = [float(np.float32(((i * 7919) % 1000) * 0.37 - 150.0)) for i in range(1000)]
    forward = sum(oracle.fixed_to_raw(x) for x in values)
    backward = sum(oracle.fixed_to_raw(x) for x in reversed(values))
    assert forward == backward


def test_fp64_fixed_point_is_round_ties_even(oracle):
    # The build's FP64 form: llrint(x * scale * 2^32), scale a power of two
    rng = np.random.default_rng(5)
    for scale in (2.0 ** -1, 1.0, 2.0 ** 9):
        S = scale * ONE
        for x in rng.standard_normal(2000) * 10.0 ** rng.uniform(-8, 3, 2000):
            assert oracle.fixed_f64(float(x), S) == round(float(x) * S)
    # exact ties go to even
    assert oracle.fixed_f64(0.5 / ONE, ONE) == 0
    assert oracle.fixed_f64(1.5 / ONE, ONE) == 2
    assert oracle.fixed_f64(-2.5 / ONE, ONE) == -2


def test_spread_values_sum_kat(oracle):
    # tests/gpu.rs:89-149 on the CPU: 2^20 values, exact total in 64 bits
    values = spread_values(1 << 20)
    expected = sum(round(float(x) * ONE) for x in values)
    assert -(1 << 63) <= expected < (1 << 63)
    raws = [oracle.fixed_to_raw(float(x)) for x in values[:4096]]
    assert raws == [round(float(x) * ONE) for x in values[:4096]]
    exact = math.fsum(float(x) for x in values)
    assert abs(expected / ONE - exact) < 1e-3


@pytest.mark.parametrize("m,expected", [(0.0, 1.0), (float("inf"), 1.0), (1.0, 2.0 ** 20), (2.0 ** 20, 1.0),
                                        (2.0 ** 20 + 1.0, 0.5), (1.5 * 0.722 * 1048576, 0.5), (3.0, 2.0 ** 18)])
def test_scale_for_closed_form(oracle, m, expected):
    # lennard_jones.rs:435-440: 2^floor(20 - log2 m), 1 for m <= 0 or not finite
    assert oracle.scale_for(m) == expected
