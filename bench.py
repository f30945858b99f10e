"""Benchmark: LJ atom-steps/s at N = 1,048,576, rho = 0.8442 on one MI355X.

Workload: FCC 4 x 64^3 at rho 0.8442, T 0.722 (Maxwell-Boltzmann, zero
momentum), rc 2.5 truncated and shifted, Verlet lists with skin 0.3 rebuilt
at skin/2, velocity Verlet dt 0.005, NVE; int64 fixed-point energy and virial.
Timed: K steps of the plain FP64 form (the headline `value`), then K steps of
the FP32-pair-force form from the same initial state. A step includes the
list-rebuild decision and every rebuild that falls in the window.

Also reported: the force kernel's roofline from HIP events on the library's
stream, and the oracle's f64 CPU path on the same workload over a bounded
sample (the CPU baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "virial-engine_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one force-kernel launch per atom: read x_i and v_i,
# write a_i and v_i, 24 bytes each in FP64.
FORCE_BYTES_PER_ATOM = {0: 96.0, 1: 96.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_form(lj, particles, side, precision, steps, warmup, timed_steps):
    settings = lj.Settings([side] * 3, 2.5, 0.005, skin=0.3, precision=precision)
    sim = lj.LennardJones(particles, settings)
    n = sim.count()
    sim.run(warmup)
    start = sim.measure_now()
    t0 = time.perf_counter()
    sim.run(steps)  # returns after the stream is synchronised
    t1 = time.perf_counter()
    end = sim.measure_now()
    rebuilds = end.list_rebuilds - start.list_rebuilds
    force_ms, total_ms = sim.timed_run(timed_steps)
    pairs = end.pairs_within_cutoff
    sim.close()
    return {
        "n": n,
        "seconds": t1 - t0,
        "atom_steps_per_s": n * steps / (t1 - t0),
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "rebuilds": rebuilds,
        "force_kernel_us": force_ms * 1e3 / timed_steps,
        "timed_ms_per_step": total_ms / timed_steps,
        "pairs_within_cutoff": pairs,
        "energy_start": start.energy(),
        "energy_end": end.energy(),
        "list_overflow": end.list_overflow,
    }


def cpu_baseline(n_side, density, temperature, seed, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = O.max_threads()
    pos, vel, side = O.fcc(n_side, density, temperature, seed)
    sim = O.Sim(pos, vel, [side] * 3, cutoff=2.5, skin=0.3, dt=0.005)
    sim.run(2)  # first lists and forces, then steady stepping
    t0 = time.perf_counter()
    sim.run(steps)
    t1 = time.perf_counter()
    n = pos.shape[0]
    return {
        "value": n * steps / (t1 - t0),
        "unit": "atom-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle f64 cell-list + Verlet-list path (OpenMP over atoms), same FCC 4x{n_side}^3 "
                  f"workload, {steps} steps after 2 warm-up steps, {t1 - t0:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=64, help="FCC cells per side (4 n^3 atoms)")
    ap.add_argument("--timed-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path has no multi-GPU form; "
              "see DESIGN.md for how it would shard", file=sys.stderr)
        sys.exit(2)

    from virial_particles import lennard_jones as lj

    density, temperature, seed = 0.8442, 0.722, 2024
    particles, side = lj.fcc_array(args.cells, density, temperature, seed)
    n = particles.shape[0]
    log(f"workload: FCC 4x{args.cells}^3 = {n} atoms, box {side}")
    timed = min(args.timed_steps, max(args.steps, 1))
    fp64 = run_form(lj, particles, side, 0, args.steps, args.warmup, timed)
    log(f"fp64: {json.dumps(fp64)}")
    fp32 = run_form(lj, particles, side, 1, args.steps, args.warmup, timed)
    log(f"fp32 forces: {json.dumps(fp32)}")

    achieved = FORCE_BYTES_PER_ATOM[0] * n / (fp64["force_kernel_us"] * 1e-6) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": None,
        "kernel": "k_force<0,false> (FP64 pair forces + second half kick)",
        "avg_launch_us": fp64["force_kernel_us"],
    }
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cells, density, temperature, seed, args.cpu_steps)
        log(f"cpu baseline: {json.dumps(cpu)}")

    line = {
        "metric": "LJ atom-steps/s, N=1M rho=0.8442: plain vs mixed-precision; batched replicas",
        "value": fp64["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"LJ NVE FCC 4x{args.cells}^3 N={n} rho={density} T={temperature} rc=2.5 skin=0.3 "
                        f"dt=0.005, FP64 state and forces",
            "n_atoms": n,
            "form": "plain FP64",
        },
        "fp32_forces": {
            "value": fp32["atom_steps_per_s"],
            "ms_per_step": fp32["ms_per_step"],
            "speedup_vs_fp64": fp32["atom_steps_per_s"] / fp64["atom_steps_per_s"],
            "force_kernel_us": fp32["force_kernel_us"],
        },
        "pairs_per_s": fp64["pairs_within_cutoff"] * args.steps / fp64["seconds"],
        "rebuilds_in_window": fp64["rebuilds"],
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
