"""Benchmark: LJ atom-steps/s at N = 1,048,576, rho = 0.8442 on one MI355X.

Workload: FCC 4 x 64^3 at rho 0.8442, T 0.722 (Maxwell-Boltzmann, zero
momentum), rc 2.5 truncated and shifted, Verlet lists with skin 0.3 rebuilt
at skin/2, velocity Verlet dt 0.005, NVE; int64 fixed-point energy and virial.
Timed: K steps of the plain FP64 form (the headline `value`), then K steps of
the FP32-pair-force form from the same initial state. A step includes the
rebuild decision and every list rebuild that falls in the window. Then the
batched form: 512 Langevin replicas of mixed size, density and temperature
in the same launches, 5000 steps, with each replica's PV/NkT and its
standard error from block means.

Also reported:
  - the force kernel's roofline: algorithmic bytes per launch over its
    average duration from HIP events on the library's stream;
  - the bandwidth a plain streaming copy reaches on the device;
  - the oracle's f64 CPU path on the same workload over a bounded sample
    (all host threads, and one thread), the CPU baseline.

Usage:
  python bench.py [--gpus 1] [--steps K] [--warmup W]
  python bench.py --precision-study --cells 128 --steps 10000  (FP32 vs FP64 forms, 8M atoms)
  python bench.py --slab --steps 2000                          (8 seeds of the 2M-atom slab)
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
# Algorithmic bytes of one force-kernel launch per atom in the NVE form,
# where the launch also drifts the next step's positions: read x_i, v_i, a_i
# and x_i at the list build; write a_i, v_i and the next x_i; 24 bytes each
# in FP64.
FORCE_BYTES_PER_ATOM = 168.0
# Flops of one pair-force evaluation (displacement 3, r^2 5, 1/r^2 1, s^6 2,
# s^12 1, r.F 3, F/r^2 1, accumulate 6) and MI355X's published FP64 vector peak.
PAIR_FLOPS = 22.0
FP64_PEAK_TFLOPS = 78.6
FORCE_KERNEL = "k_force<Layout<512, 3200, false>, 0, false, false>"
COPY_KERNEL = "k_copy"
COPY_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_force_fcc1m_traffic.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_form(lj, particles, side, precision, steps, warmup, timed_steps, long_steps=0, melt=0):
    """Times `steps` steps of one form after `melt` steps, the brick
    planner's tuning steps and `warmup` steps; then the force kernel's
    HIP-event time over `timed_steps`; then, with long_steps, the rate over
    that many more steps (a window as long as the batched form's). The plan
    is tuned after the melt: on the starting lattice the fills 0.9 and 1.0
    time within 1 us of each other, while in the liquid 1.0 ran the force
    launch ~5% slower (133.7 against 127.7 us, mean of 4 and 12 runs)."""
    settings = lj.Settings([side] * 3, 2.5, 0.005, skin=0.3, precision=precision)
    sim = lj.LennardJones(particles, settings)
    n = sim.count()
    sim.run(melt)
    tune_steps = sim.tune()
    sim.run(warmup)
    start = sim.measure_now()
    t0 = time.perf_counter()
    sim.run(steps)  # returns after the stream is synchronised
    t1 = time.perf_counter()
    end = sim.measure_now()
    force_ms, total_ms = sim.timed_run(timed_steps)
    plan = sim.plan_info()
    long = None
    if long_steps:
        c0 = sim.measure_now().list_rebuilds
        t2 = time.perf_counter()
        sim.run(long_steps)
        t3 = time.perf_counter()
        long = {"steps": long_steps, "atom_steps_per_s": n * long_steps / (t3 - t2),
                "rebuilds": sim.measure_now().list_rebuilds - c0}
    sim.close()
    return {
        "melt_steps": melt,
        "tune_steps": tune_steps,
        "plan": plan,
        "long_window": long,
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


def run_batch(lj, specs, meta, steps, warmup, block_steps, blocks):
    """Runs the batch; times the stepping only. Pressure is sampled every
    block_steps steps and PV/NkT gets a standard error from `blocks` block
    means (tests/thermodynamics.rs:259-277)."""
    import numpy as np

    b = lj.Batch(specs)
    tune_steps = b.tune()
    b.run(warmup)
    samples = []
    stepping = 0.0
    atoms = sum(m["n"] for m in meta)
    segment, seg_time, seg_steps = [], 0.0, 0  # the rate over the run, in stretches of ~500 steps
    for _ in range(max(steps // block_steps, 1)):
        t0 = time.perf_counter()
        b.run(block_steps)
        dt = time.perf_counter() - t0
        stepping += dt
        seg_time += dt
        seg_steps += block_steps
        ms = b.measure_now()
        if seg_steps >= 500:
            c = b.debug_counters()
            pairs = sum(m.pairs_within_cutoff for m in ms)  # unique pairs within rc now
            segment.append({"steps": seg_steps, "atom_steps_per_s": atoms * seg_steps / seg_time,
                            "pairs_per_s": pairs * seg_steps / seg_time, "pairs_per_atom": 2.0 * pairs / atoms,
                            "cut_builds": c["stage_overflow_builds"], "rebuilds": c["rebuilds"]})
            seg_time, seg_steps = 0.0, 0
        samples.append([m.pressure() * m.volume / (m.count * meta[r]["T"]) for r, m in enumerate(ms)])
    last = b.measure_now()
    counters = b.debug_counters()
    plan = b.plan_info()
    b.close()
    z = np.array(samples)  # samples x replicas
    usable = (z.shape[0] // blocks) * blocks
    means = z[:usable].reshape(blocks, -1, z.shape[1]).mean(axis=1)
    zmean = means.mean(axis=0)
    zse = means.std(axis=0, ddof=1) / np.sqrt(blocks)
    done = max(steps // block_steps, 1) * block_steps
    return {
        "replicas": len(specs),
        "atoms": atoms,
        "tune_steps": tune_steps,
        "plan": plan,
        "steps": done,
        "seconds_stepping": stepping,
        "atom_steps_per_s": atoms * done / stepping,
        "pairs_per_s": sum(m.pairs_within_cutoff for m in last) * done / stepping,
        "pv_nkt": [{"n": m["n"], "rho": m["rho"], "T": m["T"], "z": float(zmean[r]), "se": float(zse[r])}
                   for r, m in enumerate(meta)],
        "list_overflow": int(max(m.list_overflow for m in last)),
        "counters": counters,
        "segments": segment,
    }


def precision_study(lj, cells, steps, every, equilibrate=1000):
    """The FP32-pair-force form against the FP64 form (bit-exact with the f64
    oracle) on FCC 4 x cells^3 at rho 0.8442, T 0.722, melted for
    `equilibrate` FP64 steps: per-atom force error, PE and virial error, and
    NVE energy drift of both over `steps` steps, sampled every `every`."""
    import numpy as np

    lattice, side = lj.fcc_array(cells, 0.8442, 0.722, 2024)
    n = lattice.shape[0]
    # Melt first: on the perfect lattice every net force is zero by symmetry,
    # and the first steps carry the melting transient, not drift
    melt = lj.LennardJones(lattice, lj.Settings([side] * 3, 2.5, 0.005, skin=0.3))
    melt.run(equilibrate)
    particles = melt.read_particle_array()
    melt.close()
    out = {"workload": f"FCC 4x{cells}^3 N={n} rho=0.8442 T=0.722 rc=2.5 skin=0.3 dt=0.005 NVE, both forms "
                       f"from the state after {equilibrate} FP64 steps", "n": n, "steps": steps}
    forms = {}
    for label, prec in (("fp64", 0), ("fp32", 1)):
        sim = lj.LennardJones(particles, lj.Settings([side] * 3, 2.5, 0.005, skin=0.3, precision=prec))
        m0 = sim.measure_now()
        _, _, acc, orig = sim.debug_state()
        a = np.zeros_like(acc)
        a[orig] = acc
        energies = [(0, m0.energy())]
        t0 = time.perf_counter()
        for k in range(steps // every):
            sim.run(every)
            energies.append(((k + 1) * every, sim.measure_now().energy()))
        seconds = time.perf_counter() - t0
        sim.close()
        e = np.array([x[1] for x in energies])
        forms[label] = {"m0": m0, "acc": a, "energies": energies, "seconds": seconds,
                        "drift": float((e[-1] - e[0]) / abs(e[0])),
                        "max_abs_deviation": float(np.abs(e - e[0]).max() / abs(e[0]))}
        log(f"{label}: drift {forms[label]['drift']:.3e}, max |dE|/|E0| {forms[label]['max_abs_deviation']:.3e}, "
            f"{seconds:.1f} s")
    a64, a32 = forms["fp64"]["acc"], forms["fp32"]["acc"]
    err = np.linalg.norm(a32 - a64, axis=1)
    mag = np.linalg.norm(a64, axis=1)
    m64, m32 = forms["fp64"]["m0"], forms["fp32"]["m0"]
    out.update({
        "force_rms_relative": float(np.sqrt(np.mean(err ** 2) / np.mean(mag ** 2))),
        "force_max_relative_to_rms": float(err.max() / np.sqrt(np.mean(mag ** 2))),
        "force_median_per_atom_relative": float(np.median(err / np.maximum(mag, 1e-300))),
        "pe_relative": float((m32.potential - m64.potential) / abs(m64.potential)),
        "virial_relative": float((m32.virial - m64.virial) / abs(m64.virial)),
        "drift_fp64": forms["fp64"]["drift"],
        "drift_fp32": forms["fp32"]["drift"],
        "max_deviation_fp64": forms["fp64"]["max_abs_deviation"],
        "max_deviation_fp32": forms["fp32"]["max_abs_deviation"],
        "energy_fp64": forms["fp64"]["energies"][:: max(1, len(forms["fp64"]["energies"]) // 10)],
        "energy_fp32": forms["fp32"]["energies"][:: max(1, len(forms["fp32"]["energies"]) // 10)],
        "atom_steps_per_s_fp64": n * steps / forms["fp64"]["seconds"],
        "atom_steps_per_s_fp32": n * steps / forms["fp32"]["seconds"],
    })
    return out


def run_slab(lj, cells, seeds, steps, warmup, sample_every, bins):
    """8 seeds of the slab in one batch (force-shifted rc 2.5, Langevin gamma
    0.5 at T 0.7, FP32 pair forces, `bins` z slabs); times the stepping and
    samples the profiles every `sample_every` steps."""
    import numpy as np

    from virial_particles import workloads as W

    specs = []
    for k in range(seeds):
        parts, box = W.slab_replica(cells, 100 + 2 * k)
        specs.append((parts, W.slab_settings(box, 1000 + k, bins)))
    atoms = sum(p.shape[0] for p, _ in specs)
    b = lj.Batch(specs)
    tune_steps = b.tune()
    b.run(warmup)
    stepping = 0.0
    done = 0
    while done < steps:
        k = min(sample_every, steps - done)
        t0 = time.perf_counter()
        b.run(k)
        stepping += time.perf_counter() - t0
        b.sample_profile()
        done += k
    last = b.measure_now()
    out = {"replicas": seeds, "atoms": atoms, "steps": steps, "seconds_stepping": stepping,
           "counters": b.debug_counters(),
           "atom_steps_per_s": atoms * steps / stepping,
           "pairs_per_s": sum(m.pairs_within_cutoff for m in last) * steps / stepping}
    tensions, rho_l, rho_v = [], [], []
    for r in range(seeds):
        raw, samples = b.read_profile_raw(r)
        sums = raw / 2.0 ** 32 / samples
        prof = lj.Profile.from_sums(sums, samples, specs[r][1].box_size)
        phases = prof.phases(0.36, 4.0)
        dens = prof.density()
        rho_l.append(prof.mean_in(dens, phases, lj.Phase.LIQUID))
        rho_v.append(prof.mean_in(dens, phases, lj.Phase.VAPOUR))
        tensions.append(prof.surface_tension())
    b.close()

    def mean_se(v):
        v = np.array([x for x in v if x is not None], dtype=float)
        return (float(v.mean()), float(v.std(ddof=1) / np.sqrt(len(v)))) if len(v) > 1 else (None, None)

    out["surface_tension"] = mean_se(tensions)
    out["liquid_density"] = mean_se(rho_l)
    out["vapour_density"] = mean_se(rho_v)
    out["profile_samples"] = int(steps // sample_every)
    out["workload"] = (f"{seeds} seeds of a liquid-vapour slab, Lz = 4 Lx, FCC liquid (rho 0.725) in the middle "
                       f"quarter, vapour rho 0.02; force-shifted rc 2.5, Langevin gamma 0.5 T 0.7, FP32 pair "
                       f"forces, {bins} z slabs")
    return out


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


def working_means(path):
    """Counter means over each kernel's working dispatches: those whose busy
    cycles (or, without that counter, whose sum of all counters) reach half
    of the kernel's largest. The list build and the rebuild chain return at
    once on steps without a rebuild; their plain means mix both."""
    c = _db(path)
    per = defaultdict(lambda: defaultdict(dict))
    for did, k, cn, v in c.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        per[_short(k)][did][cn] = v
    out = {}
    for k, ds in per.items():
        key = lambda r: r.get("SQ_BUSY_CYCLES", sum(r.values()))
        top = max(key(r) for r in ds.values())
        work = [r for r in ds.values() if key(r) >= 0.5 * top]
        out[k] = (len(work), len(ds), {cn: sum(r[cn] for r in work) / len(work) for cn in work[0]})
    return out


def summarize(path, out, command):
    lines = [f"# rocprofv3 summary of: {command}", ""]
    c = _db(path)
    has_counters = c.execute("select count(*) from sqlite_master where name = 'counters_collection'").fetchone()[0]
    if has_counters and c.execute("select count(*) from counters_collection").fetchone()[0]:
        lines.append(f"{'kernel':60s} counter means per dispatch")
        for k, cs in sorted(counter_means(path).items()):
            lines.append(f"{k[:60]:60s} " + ", ".join(f"{cn}={v:.1f}" for cn, v in sorted(cs.items())))
        lines += ["", f"{'kernel':60s} counter means over the working dispatches (at least half the largest)"]
        for k, (n, of, cs) in sorted(working_means(path).items()):
            lines.append(f"{k[:60]:60s} [{n}/{of}] " + ", ".join(f"{cn}={v:.1f}" for cn, v in sorted(cs.items())))
    else:
        stats = kernel_stats(path)
        lines.append(f"{'kernel':60s} {'calls':>6s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'total_ms':>10s}")
        for k, v in sorted(stats.items(), key=lambda kv: -sum(kv[1])):
            lines.append(f"{k[:60]:60s} {len(v):6d} {sum(v) / len(v):10.2f} {min(v):10.2f} {max(v):10.2f} "
                         f"{sum(v) / 1e3:10.3f}")
        # The list build and the rebuild chain return at once on steps
        # without a rebuild: their working dispatches (at least half the
        # longest) give the time per rebuild.
        lines += ["", f"{'kernel':60s} {'working':>8s} {'avg_us':>10s}   (dispatches of at least half the longest)"]
        for k, v in sorted(stats.items(), key=lambda kv: -sum(kv[1])):
            work = [t for t in v if t >= 0.5 * max(v)]
            lines.append(f"{k[:60]:60s} {len(work):4d}/{len(v):<4d} {sum(work) / len(work):10.2f}")
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
    ap.add_argument("--melt", type=int, default=200,
                    help="untimed steps of the lattice melting before the brick plan is tuned")
    ap.add_argument("--cells", type=int, default=64, help="FCC cells per side (4 n^3 atoms)")
    ap.add_argument("--timed-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--cpu-single-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision-study", action="store_true",
                    help="FP32 vs FP64 forms on FCC 4 x cells^3 (use --cells 128 --steps 10000)")
    ap.add_argument("--measure-every", type=int, default=500)
    ap.add_argument("--slab", action="store_true", help="8 seeds of the 2M-atom liquid-vapour slab")
    ap.add_argument("--slab-cells", type=int, default=77, help="FCC cells per side of the slab's liquid block")
    ap.add_argument("--batch-replicas", type=int, default=512, help="0 skips the batched form")
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--batch-table", help="write each replica's PV/NkT and standard error here (JSON)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched form")
    ap.add_argument("--physics", action="store_true",
                    help="the reference's physical validations, full protocols (JSON per check)")
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
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): this path has no multi-GPU form; "
              "replicas would be split across GPUs by cost, a single system into z slabs", file=sys.stderr)
        sys.exit(2)

    from virial_particles import lennard_jones as lj
    from virial_particles import workloads as W

    if args.physics:
        from virial_particles import validation as V

        checks = [("mb_relaxation", V.mb_relaxation, {}),
                  ("dt_study", V.dt_study, {}),
                  ("coexistence_force_shifted", V.coexistence, {"shifted_force": True, "blocks": 5, "equilibrate": 60_000}),
                  ("coexistence_truncated", V.coexistence, {"shifted_force": False, "blocks": 5, "equilibrate": 60_000}),
                  ("fourier_green_kubo", V.fourier_green_kubo, {})]
        for name, f, kw in checks:
            t0 = time.perf_counter()
            r = f(**kw)
            r["seconds"] = time.perf_counter() - t0
            print(json.dumps({"check": name, **r}, default=str), flush=True)
        return
    if args.slab:
        print(json.dumps(run_slab(lj, args.slab_cells, 8, args.steps, args.warmup, 100, 200)), flush=True)
        return
    if args.precision_study:
        print(json.dumps(precision_study(lj, args.cells, args.steps, args.measure_every)), flush=True)
        return
    if args.batch_only:
        specs, meta = W.batch_workload(args.batch_replicas, 7)
        br = run_batch(lj, specs, meta, args.batch_steps, args.warmup, 25, 20)
        br.pop("pv_nkt")
        print(json.dumps(br), flush=True)
        return

    copy_ms = lj.copy_bandwidth(COPY_BYTES, 10)
    copy_gbs = 2.0 * COPY_BYTES / (copy_ms * 1e-3) / 1e9
    log(f"plain copy of {COPY_BYTES >> 20} MiB: {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s (read + write)")
    fma_tflops = lj.fp64_fma_rate(10)
    log(f"FP64 FMA chains: {fma_tflops:.1f} TFLOP/s")

    density, temperature, seed = 0.8442, 0.722, 2024
    particles, side = lj.fcc_array(args.cells, density, temperature, seed)
    n = particles.shape[0]
    log(f"workload: FCC 4x{args.cells}^3 = {n} atoms, box {side}")
    timed = min(args.timed_steps, max(args.steps, 1))
    fp64 = run_form(lj, particles, side, 0, args.steps, args.warmup, timed,
                    long_steps=args.batch_steps if args.batch_replicas > 0 else 0, melt=args.melt)
    log(f"fp64: {json.dumps(fp64)}")
    fp32 = run_form(lj, particles, side, 1, args.steps, args.warmup, timed, melt=args.melt)
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
        "kernel": FORCE_KERNEL + ": FP64 pair forces from brick-staged lists, the velocity step and the next step's drift",
        "algorithmic_bytes_per_launch": FORCE_BYTES_PER_ATOM * n,
        "avg_launch_us": fp64["force_kernel_us"],
        "measured_copy_GBs": copy_gbs,
    }
    # The same kernel against the FP64 vector peak: 22 flops per unique pair
    # within rc (P_u from the pair counter) per atom-step; a full list
    # evaluates each pair twice, which this count does not credit.
    p_u = fp64["pairs_within_cutoff"] / n
    valu_achieved = PAIR_FLOPS * p_u * n / (fp64["force_kernel_us"] * 1e-6) / 1e12
    valu = {
        "bound": "fp64 valu",
        "achieved": valu_achieved,
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": valu_achieved / FP64_PEAK_TFLOPS,
        "flops_per_atom_step": PAIR_FLOPS * p_u,
        "measured_fma_TFLOPs": fma_tflops,
        "frac_of_measured": valu_achieved / fma_tflops if fma_tflops > 0 else None,
    }
    batched = None
    if args.batch_replicas > 0:
        specs, meta = W.batch_workload(args.batch_replicas, 7)
        br = run_batch(lj, specs, meta, args.batch_steps, args.warmup, 25, 20)
        zs = br.pop("pv_nkt")
        if args.batch_table:
            json.dump(zs, open(args.batch_table, "w"), indent=0)
        se_rel = sorted(z["se"] / abs(z["z"]) for z in zs if z["z"] != 0)
        batched = dict(br)
        # against the single FP64 system over a window as long as the batch's
        # (the same run, after the headline window), not the short headline window
        batched["ratio_vs_single_fp64"] = br["atom_steps_per_s"] / fp64["long_window"]["atom_steps_per_s"]
        batched["single_fp64_same_window"] = fp64["long_window"]
        batched["pv_nkt_median_relative_se"] = se_rel[len(se_rel) // 2] if se_rel else None
        batched["workload"] = (f"{args.batch_replicas} Langevin replicas (gamma 1, T 0.7-2.0), N 500-50000 "
                               f"log-uniform, rho 0.01-0.85, FCC or random insertion, rc 2.5, dt 0.005, FP64; "
                               f"pressure every 25 steps, 20 blocks")
        log(f"batched: {json.dumps(batched)}")
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
            "target_speedup": 2.0,
            "target_met": fp32["atom_steps_per_s"] >= 2.0 * fp64["atom_steps_per_s"],
            "force_kernel_us": fp32["force_kernel_us"],
        },
        "pairs_per_s": fp64["pairs_within_cutoff"] * args.steps / fp64["seconds"],
        "rebuilds_in_window": fp64["rebuilds"],
        "timed_from_step": fp64["melt_steps"] + fp64["tune_steps"] + args.warmup,
        "plan": {"fill": fp64["plan"]["fill"], "trials": fp64["plan"]["trials"]},
        # NVE from the lattice, the window from step timed_from_step: the
        # relative energy change over it holds what is left of the melting
        # transient, not only a drift (the 10^4-step drift is measured after
        # melting: --precision-study)
        "energy_change_in_window": (fp64["energy_end"] - fp64["energy_start"]) / abs(fp64["energy_start"]),
        "batched": batched,
        "roofline": roofline,
        "valu_roofline": valu,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
