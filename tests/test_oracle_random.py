"""Counter-based noise of the Langevin thermostat: Philox4x32-10 against the
published known-answer vectors (Random123, Salmon et al. SC'11), and the
basic-operation log / sin / cos against libm."""
import math

import numpy as np
import pytest

KATS = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expected", KATS)
def test_philox_known_answers(oracle, ctr, key, expected):
    assert list(oracle.philox4x32(ctr, key)) == expected


def test_log_against_libm(oracle):
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(0, 1, 20000), 10.0 ** rng.uniform(-300, 300, 5000),
                         [2.3283064365386963e-10 * 0.5, 1.0 - 2 ** -53, 0.5, 1.0, 2.0, math.sqrt(2.0)]])
    for x in xs:
        if x <= 0:
            continue
        got, want = oracle.log(float(x)), math.log(float(x))
        assert abs(got - want) <= 4.5e-16 * max(abs(want), 1e-300) + 1e-300, x
    assert oracle.log(1.0) == 0.0


def test_sincos_turns_against_libm(oracle):
    rng = np.random.default_rng(4)
    for u in np.concatenate([rng.uniform(0, 1, 20000), np.arange(0, 1.0001, 1 / 64)]):
        s, c = oracle.sincos_turns(float(u))
        assert abs(s - math.sin(2 * math.pi * u)) < 1e-15
        assert abs(c - math.cos(2 * math.pi * u)) < 1e-15
    assert oracle.sincos_turns(0.0) == (0.0, 1.0)
    assert oracle.sincos_turns(0.25)[0] == 1.0
    assert oracle.sincos_turns(0.5)[1] == -1.0


def test_noise_is_standard_normal_and_keyed(oracle):
    draws = np.array([oracle.langevin_noise(7, 0, i, s) for i in range(200) for s in range(100)]).ravel()
    assert abs(draws.mean()) < 0.02
    assert abs(draws.var() - 1.0) < 0.03
    kurt = np.mean(draws ** 4) / np.mean(draws ** 2) ** 2
    assert abs(kurt - 3.0) < 0.15
    # same key, same numbers; any key change, different numbers
    a = oracle.langevin_noise(7, 0, 5, 11)
    assert np.array_equal(a, oracle.langevin_noise(7, 0, 5, 11))
    for other in [(8, 0, 5, 11), (7, 1, 5, 11), (7, 0, 6, 11), (7, 0, 5, 12)]:
        assert not np.array_equal(a, oracle.langevin_noise(*other))
