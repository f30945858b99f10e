"""Writes tests/golden/state_two_replicas.vljstate: the state file of a
small batch (two liquid replicas, 30 Langevin steps), for the CPU tests of
vlj_load's checks. Needs the GPU library; run from the repository root:
    python tests/golden/make_state_fixture.py [out_path]"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "virial-engine_amd"))

from virial_particles import lennard_jones as lj  # noqa: E402


def main(out):
    specs = []
    for k, n in enumerate((216, 300)):
        parts, side = lj.fcc_count_array(n, 0.8, 1.0, 31 + k)
        specs.append((parts, lj.Settings([side] * 3, 2.5, 0.005, skin=0.3, thermostat=lj.Langevin(1.0, 1.0, 5 + k))))
    b = lj.Batch(specs)
    b.run(30)
    b.save(out)
    b.close()
    print(f"wrote {out} ({os.path.getsize(out)} bytes)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(ROOT, "tests", "golden", "state_two_replicas.vljstate"))
