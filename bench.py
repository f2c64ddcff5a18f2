"""Benchmark of the SPH step on one MI355X.

Workload (the metric's configuration): the 1,000,000-particle dam break
(100^3 jittered lattice, spacing 0.5 h, h = 0.03, dt = 0.001, mass 1.6e-4,
box [-2,2] x [-1,1]^2), state resident in HBM.  A step is one pass of the SPH
operator over all particles (scan, scatter, reorder, density, force +
integrate).  The timed job is the workload's first K steps (default its full
1000) from the initial state, replayed as HIP graphs after W untimed warm-up
steps on the same context (re-uploaded before the timed run); `value` is
particle-updates/s = N * K / wall time, bracketed by device synchronisation.

The JSON line also carries
  roofline      the dominant kernel's algorithmic bytes per launch (the
                per-particle figure of its passes times N) over its mean
                duration, timed with HIP events on the kernel's stream over
                the same K-step job, against the 8 TB/s HBM peak; `traffic`
                is the measured HBM bytes per launch when a counter summary
                for the kernel is present under profiles/
  cpu_baseline  the oracle (CPU restatement of the reference's
                multithreaded update()) timed on this host over a few steps
                of the same workload
"""
import argparse
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "sph-fluid-simulator_amd", "python"))

import numpy as np  # noqa: E402

import sph_amd  # noqa: E402
import sph_scenes  # noqa: E402

METRIC = "particle-updates/s at 1M (1 GPU, % HBM roofline); batched scene-steps/s + p99"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle and step of each kernel: the per-pass
# figures (reads + writes, neighbour reads counted once) of the step's
# 220-byte budget: cell ranges 4, sort 16, reorder gather 52, density 20,
# force 44 + integrate 64 + next cell key 20.
ALG_BYTES = {"scan": 4, "scatter": 16, "reorder": 52, "density": 20, "force_integrate": 128}
ALG_BYTES_STEP = 220


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return min(os.cpu_count() or 1, 16)


def cpu_baseline(sc, budget_s=15.0, max_steps=20):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_binding import Oracle

    orc = Oracle(32)
    s = orc.settings(*sc.settings_args())
    box = orc.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    p = orc.particles(sc.pos, sc.vel)
    threads = cpu_threads()
    times = []
    t_start = time.perf_counter()
    while len(times) < max_steps and time.perf_counter() - t_start < budget_s:
        t0 = time.perf_counter()
        orc.update(p, s, sc.dt, box, nthreads=threads, transforms=False)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {
        "value": sc.n / med,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(times)} steps of the {sc.n}-particle workload from its initial state "
                  f"(median {med * 1e3:.1f} ms/step, {threads} threads, {platform.processor() or platform.machine()})",
    }


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from a committed counter summary
    (profiles/traffic_<kernel>_workload2.json), or None."""
    path = os.path.join(REPO, "profiles", f"traffic_{kernel}_workload2.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel event pass")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run the copy kernel over this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path", file=sys.stderr)
        sys.exit(2)

    sc = sph_scenes.workload(args.workload)
    log(f"workload {sc.name}: n={sc.n} h={sc.h} dt={sc.dt}")
    s = sph_amd.settings(*sc.settings_args())
    b = sph_amd.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    ctx = sph_amd.Context(sc.n, s, b, sc.dt)
    ctx.upload(sc.pos, sc.vel)
    ctx.step(args.warmup)
    ctx.sync()
    ctx.upload(sc.pos, sc.vel)  # the timed job starts from the initial state
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = sc.n * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step, {value / 1e9:.3f} G particle-updates/s")

    roofline = None
    if not args.no_profile:
        ctx.upload(sc.pos, sc.vel)  # same job again, with events around every kernel
        prof = ctx.profile(args.steps)
        log("per-kernel mean ms (HIP events): " + ", ".join(f"{k}={v:.4f}" for k, v in prof.items()))
        dom = max(prof, key=prof.get)
        ach = ALG_BYTES[dom] * sc.n / (prof[dom] * 1e-3) / 1e9
        roofline = {
            "bound": "hbm",
            "achieved": round(ach, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(dom),
            "kernel": dom,
            "kernel_ms": round(prof[dom], 5),
            "alg_bytes_per_particle": ALG_BYTES[dom],
            "step_alg_GBs": round(ALG_BYTES_STEP * value / 1e9, 2),
            "step_frac": round(ALG_BYTES_STEP * value / 1e9 / HBM_PEAK_GBS, 4),
            "per_kernel_ms": {k: round(v, 5) for k, v in prof.items()},
        }
    ctx.close()

    if args.copy_bytes > 0:
        ms = sph_amd.copy_bench(args.copy_bytes, 10)
        log(f"copy kernel: {args.copy_bytes} B in {ms:.4f} ms = {2 * args.copy_bytes / ms / 1e6:.1f} GB/s (read+write)")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, multithreaded) ...")
        cpu = cpu_baseline(sc)
        log(f"cpu baseline: {cpu['value'] / 1e6:.3f} M particle-updates/s")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{sc.name}: {sc.n} particles, 100^3 jittered lattice dam break, h={sc.h}, "
                        f"dt={sc.dt}, box [-2,2]x[-1,1]^2, steps 1..{args.steps} of the job, resident in HBM",
            "particles": sc.n,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
