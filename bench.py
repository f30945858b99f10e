"""Benchmark: LJ atom-steps/s at N = 1,048,576, rho = 0.8442 on one MI355X.

Workload: FCC 4 x 64^3 at rho 0.8442, T 0.722 (Maxwell-Boltzmann, zero
momentum), rc 2.5 truncated and shifted, Verlet lists with skin 0.3 rebuilt
at skin/2, velocity Verlet dt 0.005, NVE; int64 fixed-point energy and virial.
Timed: K steps of the plain FP64 form (the headline `value`), then K steps of
the FP32-pair-force form from the same initial state. A step includes the
rebuild decision and every list rebuild that falls in the window.

Also reported:
  - the force kernel's roofline: algorithmic bytes per launch over its
    average duration from HIP events on the library's stream;
  - the bandwidth a plain streaming copy reaches on the device;
  - the oracle's f64 CPU path on the same workload over a bounded sample
    (all host threads, and one thread), the CPU baseline.

Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W]
  python bench.py --summarize ROCPROF_DIR --out FILE     (text summary of a rocprofv3 run)
  python bench.py --traffic FETCH_DIR WRITE_DIR --out FILE (calibrated bytes per force launch)
"""
import argparse
import glob
import json
import os
import sqlite3
import sys
import time
from collections import defaultdict

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "virial-engine_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
# Algorithmic bytes of one force-kernel launch per atom: read x_i and v_i,
# write a_i and v_i, 24 bytes each in FP64.
FORCE_BYTES_PER_ATOM = 96.0
FORCE_KERNEL = "k_force<0, false>"
COPY_KERNEL = "k_copy"
COPY_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_force_fcc1m_traffic.txt")


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
    force_ms, total_ms = sim.timed_run(timed_steps)
    sim.close()
    return {
        "n": n,
        "seconds": t1 - t0,
        "atom_steps_per_s": n * steps / (t1 - t0),
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "rebuilds": end.list_rebuilds - start.list_rebuilds,
        "force_kernel_us": force_ms * 1e3 / timed_steps,
        "timed_ms_per_step": total_ms / timed_steps,
        "pairs_within_cutoff": end.pairs_within_cutoff,
        "energy_start": start.energy(),
        "energy_end": end.energy(),
        "list_overflow": end.list_overflow,
    }


def cpu_baseline(n_side, density, temperature, seed, steps, single_steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = O.max_threads()
    pos, vel, side = O.fcc(n_side, density, temperature, seed)
    n = pos.shape[0]
    out = {}
    for label, th, k in (("all", threads, steps), ("one", 1, single_steps)):
        O.set_threads(th)
        sim = O.Sim(pos, vel, [side] * 3, cutoff=2.5, skin=0.3, dt=0.005)
        sim.run(2)  # first lists and forces, then steady stepping
        t0 = time.perf_counter()
        sim.run(k)
        t1 = time.perf_counter()
        out[label] = (n * k / (t1 - t0), t1 - t0, k)
        sim.close()
    O.set_threads(threads)
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": out["all"][0],
        "unit": "atom-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle f64 cell-list + Verlet-list path (OpenMP over atoms, {threads} threads on {cpu}), "
                  f"same FCC 4x{n_side}^3 workload, {out['all'][2]} steps after 2 warm-up steps ({out['all'][1]:.1f} s, "
                  f"list rebuilds included)",
        "single_thread": {"value": out["one"][0], "steps": out["one"][2], "seconds": out["one"][1]},
    }


# ---- rocprofv3 database summaries ----

def _db(path):
    dbs = glob.glob(os.path.join(path, "**", "*.db"), recursive=True)
    if not dbs:
        raise SystemExit(f"no rocprofv3 database under {path}")
    return sqlite3.connect(dbs[0])


def _short(name):
    return name.replace("vlj::", "").replace("(vlj::Buffers)", "").replace("(vlj::Buffers, int)", "").replace(
        "void ", "").strip()


def kernel_stats(path):
    c = _db(path)
    rows = c.execute("select name, end - start from kernels").fetchall()
    d = defaultdict(list)
    for name, dur in rows:
        d[_short(name)].append(dur / 1000.0)
    return d


def counter_means(path):
    c = _db(path)
    d = defaultdict(lambda: defaultdict(list))
    for k, cn, v in c.execute("select kernel_name, counter_name, value from counters_collection"):
        d[_short(k)][cn].append(v)
    return {k: {cn: sum(v) / len(v) for cn, v in cs.items()} for k, cs in d.items()}


def summarize(path, out, command):
    lines = [f"# rocprofv3 summary of: {command}", ""]
    c = _db(path)
    has_counters = c.execute("select count(*) from sqlite_master where name = 'counters_collection'").fetchone()[0]
    if has_counters and c.execute("select count(*) from counters_collection").fetchone()[0]:
        lines.append(f"{'kernel':44s} counter means per dispatch")
        for k, cs in sorted(counter_means(path).items()):
            lines.append(f"{k[:44]:44s} " + ", ".join(f"{cn}={v:.1f}" for cn, v in sorted(cs.items())))
    else:
        stats = kernel_stats(path)
        lines.append(f"{'kernel':44s} {'calls':>6s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'total_ms':>10s}")
        for k, v in sorted(stats.items(), key=lambda kv: -sum(kv[1])):
            lines.append(f"{k[:44]:44s} {len(v):6d} {sum(v) / len(v):10.2f} {min(v):10.2f} {max(v):10.2f} "
                         f"{sum(v) / 1e3:10.3f}")
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def traffic(fetch_dir, write_dir, out, command):
    """HBM bytes per force launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
    scaled by what the same counters read for the plain copy of COPY_BYTES."""
    f = counter_means(fetch_dir)
    w = counter_means(write_dir)
    copy = [k for k in f if COPY_KERNEL in k][0]
    force = [k for k in f if FORCE_KERNEL in k][0]
    fetch_scale = COPY_BYTES / (f[copy]["FETCH_SIZE"] * 1024.0)
    write_scale = COPY_BYTES / (w[copy]["WRITE_SIZE"] * 1024.0)
    rec = {
        "command": command,
        "force_kernel": force,
        "fetch_kib_raw": f[force]["FETCH_SIZE"],
        "write_kib_raw": w[force]["WRITE_SIZE"],
        "fetch_scale": fetch_scale,
        "write_scale": write_scale,
        "fetch_bytes": f[force]["FETCH_SIZE"] * 1024.0 * fetch_scale,
        "write_bytes": w[force]["WRITE_SIZE"] * 1024.0 * write_scale,
    }
    rec["bytes_per_launch"] = rec["fetch_bytes"] + rec["write_bytes"]
    json.dump(rec, open(out, "w"), indent=1)
    print(json.dumps(rec, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=64, help="FCC cells per side (4 n^3 atoms)")
    ap.add_argument("--timed-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-single-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--summarize", metavar="ROCPROF_DIR")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--out")
    ap.add_argument("--command", default="", help="profiled command, recorded in the summary")
    args = ap.parse_args()
    if args.summarize:
        summarize(args.summarize, args.out, args.command)
        return
    if args.traffic:
        traffic(args.traffic[0], args.traffic[1], args.out, args.command)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): this path has no multi-GPU form "
              "(DESIGN.md section 7 says how it would shard)", file=sys.stderr)
        sys.exit(2)

    from virial_particles import lennard_jones as lj

    copy_ms = lj.copy_bandwidth(COPY_BYTES, 10)
    copy_gbs = 2.0 * COPY_BYTES / (copy_ms * 1e-3) / 1e9
    log(f"plain copy of {COPY_BYTES >> 20} MiB: {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s (read + write)")

    density, temperature, seed = 0.8442, 0.722, 2024
    particles, side = lj.fcc_array(args.cells, density, temperature, seed)
    n = particles.shape[0]
    log(f"workload: FCC 4x{args.cells}^3 = {n} atoms, box {side}")
    timed = min(args.timed_steps, max(args.steps, 1))
    fp64 = run_form(lj, particles, side, 0, args.steps, args.warmup, timed)
    log(f"fp64: {json.dumps(fp64)}")
    fp32 = run_form(lj, particles, side, 1, args.steps, args.warmup, timed)
    log(f"fp32 forces: {json.dumps(fp32)}")

    achieved = FORCE_BYTES_PER_ATOM * n / (fp64["force_kernel_us"] * 1e-6) / 1e9
    traffic_bytes = None
    if args.cells == 64 and os.path.exists(TRAFFIC_FILE):
        traffic_bytes = json.load(open(TRAFFIC_FILE))["bytes_per_launch"]
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic_bytes,
        "kernel": "k_force<0,false>: FP64 pair forces from brick-staged lists + second half kick",
        "algorithmic_bytes_per_launch": FORCE_BYTES_PER_ATOM * n,
        "avg_launch_us": fp64["force_kernel_us"],
        "measured_copy_GBs": copy_gbs,
    }
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cells, density, temperature, seed, args.cpu_steps, args.cpu_single_steps)
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
        "energy_drift_in_window": (fp64["energy_end"] - fp64["energy_start"]) / abs(fp64["energy_start"]),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
