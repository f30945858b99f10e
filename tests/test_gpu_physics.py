"""The reference's physical validations, run on the GPU library with the
reference's own protocols and criteria (virial_particles/validation.py
cites each test). Statistical checks: a failure here is a physics result,
not a bit-level mismatch; their thresholds are the reference's."""
import json

import pytest

from virial_particles import validation as V

pytestmark = pytest.mark.gpu


def show(name, r):
    print(f"{name}: " + json.dumps(r, default=str))


def test_gas_relaxes_to_maxwell_boltzmann_at_constant_energy():
    # thermodynamics.rs:152-214, full protocol
    r = V.mb_relaxation()
    show("MB relaxation", r)
    assert r["ks_start"] > 0.2
    assert r["ks_end"] < 0.01
    assert abs(r["drift"]) < 1e-4


def test_force_shifted_slab_coexists_at_nist_densities():
    # coexistence.rs:293-345: eight seeds of the reference's slab in one
    # batch, two blocks of 20,000 steps each after 60,000 to equilibrate
    # (the vapour takes some 100 time units to fill; the reference's ten
    # blocks after 20,000 steps span 1,000)
    r = V.coexistence(True, seeds=8, equilibrate=60_000, blocks=2, block_steps=20_000)
    show("coexistence, force-shifted", r)
    assert r["blocks_with_one_slab"] == r["blocks"]
    assert r["checks"] == {k: True for k in r["checks"]}


def test_truncated_shifted_slab_has_published_surface_tension():
    # coexistence.rs:347-381
    r = V.coexistence(False, seeds=8, equilibrate=60_000, blocks=2, block_steps=20_000)
    show("coexistence, truncated and shifted", r)
    assert r["blocks_with_one_slab"] == r["blocks"]
    assert r["checks"] == {k: True for k in r["checks"]}


def test_conductivity_from_fourier_law_agrees_with_green_kubo():
    # heat.rs:366-396 with half the blocks and shorter blocks (the full
    # protocol runs from bench.py --physics)
    r = V.fourier_green_kubo(nemd_blocks=4, nemd_steady=30_000, nemd_samples=250, gk_blocks=4, gk_samples=25_000)
    show("Fourier vs Green-Kubo", r)
    assert r["passed"]


def test_energy_drift_falls_with_the_step_size():
    # lj_dt_study.rs:9-42, full protocol
    r = V.dt_study()
    show("dt study", r)
    assert r["passed"]
