"""The plain FP64 force kernel applies the cutoff by giving r^2 of a pair
outside it the sign and exponent word 0x56700000 (2^360), keeping its low
word. These checks pin the arithmetic that makes such a pair add nothing:
1/r^2 stays a normal number, s^6 rounds to exactly zero and the force factor
is -0, so x + d * fscal returns x for every accumulator value."""
import struct

import numpy as np


def with_high_word(r2: float, hi: int) -> float:
    lo = struct.unpack("<Q", struct.pack("<d", r2))[0] & 0xFFFFFFFF
    return struct.unpack("<d", struct.pack("<Q", (hi << 32) | lo))[0]


def test_far_pair_factor_is_negative_zero():
    rng = np.random.default_rng(5)
    r2s = np.concatenate([rng.uniform(6.25, 1e3, 2000), [6.25, 7.84, 3.0e36, 1e300, 1.5e-300]])
    for r2 in r2s:
        far = with_high_word(float(r2), 0x56700000)
        assert 2.0 ** 360 <= far < 2.0 ** 361
        inv2 = 1.0 / far
        assert inv2 >= np.finfo(np.float64).tiny  # normal: the reciprocal sequence holds
        s6 = inv2 * inv2 * inv2
        assert s6 == 0.0
        fscal = s6 * (s6 * 2.0 - 1.0) * inv2
        assert fscal == 0.0 and np.signbit(fscal)
        for acc in (0.0, 1.0, -2.5e-300, 7.0e12):
            for d in (-2.7, 0.0, 2.7):
                out = acc + d * fscal
                assert out == acc and np.signbit(out) == np.signbit(acc)
