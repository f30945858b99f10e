"""Test configuration: the `gpu` marker, import paths and shared fixtures.

CPU tests (`-m "not gpu"`) pin the oracle and the host logic; GPU tests
(`-m gpu`) run the HIP library through its C ABI against the oracle.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "virial-engine_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def lj():
    from virial_particles import lennard_jones

    return lennard_jones
