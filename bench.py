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
                for the kernel is present under profiles/; `copy_kernel_GBs`
                the bandwidth a plain copy kernel reaches (read + write);
                `per_kernel_alg_GBs` every pass's algorithmic GB/s and
                `l2_hit` the per-pass L2 hit rates of the committed counter
                summary (early and late in the job)
  cpu_baseline  the oracle (CPU restatement of the reference's
                multithreaded update()) timed on this host over a few steps
                of the same workload
  batched       the batched workload: 256 independent scenes (1k-64k
                particles each, log-uniform) advanced together for 500
                steps after the autotuner picked its settings; scene-steps/s,
                per-step p50/p99, and the same scenes stepped one after
                another on the plain single-scene path for comparison
  serving       the serving loop under Poisson request streams (2k-500k
                particles, 10-1000 steps, dam-break or droplet) at several
                rates: per-step p50/p99 and aggregate particle-updates/s
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
import sph_serving  # noqa: E402

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


def cpu_model():
    """The host CPU's model name and its logical CPU count, as lscpu reports them."""
    name = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name}, {os.cpu_count()} logical CPUs visible"


def job_states(sc, steps, nphases=5):
    """States of the workload's job at evenly spaced steps (0 .. steps-1),
    taken from the GPU run: the step cost grows as the fluid clumps, so the
    CPU is timed on every phase of the job, not only its cheap start."""
    phases = sorted({round(f * (steps - 1) / (nphases - 1)) for f in range(nphases)})
    ctx = sph_amd.Context(sc.n, sph_amd.settings(*sc.settings_args()),
                          sph_amd.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt)
    ctx.upload(sc.pos, sc.vel)
    out, at = [], 0
    for ph in phases:
        ctx.step(ph - at)
        at = ph
        d = ctx.download()
        out.append((ph, d["pos"], d["vel"]))
    ctx.close()
    return out


def cpu_baseline(sc, states, budget_s=30.0):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_binding import Oracle

    orc = Oracle(32)
    s = orc.settings(*sc.settings_args())
    box = orc.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    threads = cpu_threads()
    times, used = [], []
    t_start = time.perf_counter()
    for ph, pos, vel in states:
        if times and time.perf_counter() - t_start > budget_s:
            break
        p = orc.particles(pos, vel)
        t0 = time.perf_counter()
        orc.update(p, s, sc.dt, box, nthreads=threads, transforms=False)
        times.append(time.perf_counter() - t0)
        used.append(ph)
        log(f"cpu step from the job's state at step {ph}: {times[-1]:.3f} s")
    mean = float(np.mean(times))
    return {
        "value": sc.n / mean,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one oracle step from each of the job's states at steps {used} (GPU snapshots of the "
                  f"{sc.n}-particle workload), mean {mean * 1e3:.1f} ms/step, per phase "
                  f"{[round(t * 1e3, 1) for t in times]} ms, {threads} threads on {cpu_model()}",
    }


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from a committed counter summary
    (profiles/traffic_<kernel>_workload2.json), or None."""
    path = os.path.join(REPO, "profiles", f"traffic_{kernel}_workload2.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def load_valu_busy(kernel, window):
    """VALU-busy fraction of `kernel` from a committed counter summary
    (profiles/valu_busy_workload2.json: window -> kernel -> fraction), or
    None."""
    path = os.path.join(REPO, "profiles", "valu_busy_workload2.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("valu_busy", {}).get(window, {}).get(kernel)


def roofline_of(prof, n, value, window):
    """The roofline object of a per-kernel profile: the dominant kernel's
    algorithmic bytes per launch over its mean launch time against the HBM
    peak, beside its VALU-busy fraction (the passes are bound by their
    arithmetic and load latency, not by HBM)."""
    dom = max(prof, key=prof.get)
    ach = ALG_BYTES[dom] * n / (prof[dom] * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "achieved": round(ach, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(ach / HBM_PEAK_GBS, 4),
        "traffic": load_traffic(dom),
        "kernel": dom,
        "kernel_ms": round(prof[dom], 5),
        "valu_busy": load_valu_busy(dom, window),
        "alg_bytes_per_particle": ALG_BYTES[dom],
        "step_alg_GBs": round(ALG_BYTES_STEP * value / 1e9, 2),
        "step_frac": round(ALG_BYTES_STEP * value / 1e9 / HBM_PEAK_GBS, 4),
        "per_kernel_ms": {k: round(v, 5) for k, v in prof.items()},
        # each pass's algorithmic bytes per launch over its mean duration
        "per_kernel_alg_GBs": {k: round(ALG_BYTES[k] * n / (v * 1e-3) / 1e9, 2) for k, v in prof.items()
                               if k in ALG_BYTES and v > 0},
        "l2_hit": load_l2_hit(),
    }


def load_l2_hit():
    """Per-pass L2 hit rates (early and late in the job) from a committed
    counter summary (profiles/l2_hit_workload2.json), or None."""
    path = os.path.join(REPO, "profiles", "l2_hit_workload2.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hit_rate")


def dropin_leg(sc, calls=5):
    """The stateless drop-in (host particle array in, host array and render
    transforms out, as updateParticles is called): PCIe-inclusive rate."""
    lib = sph_amd.load()
    p = sph_amd.make_particles(sc.pos, sc.vel)
    xf = np.zeros((sc.n, 16), np.float32)
    s = sph_amd.settings(*sc.settings_args())
    b = sph_amd.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    import ctypes

    def call():
        rc = lib.sph_hip_update_particles_box(p.ctypes.data_as(ctypes.c_void_p), xf.ctypes.data_as(ctypes.c_void_p),
                                              sc.n, ctypes.byref(s), sc.dt, ctypes.byref(b))
        if rc != 0:
            raise sph_amd.SphError(rc, lib.sph_hip_last_error().decode())

    call()  # device buffers are allocated on the first call
    t0 = time.perf_counter()
    for _ in range(calls):
        call()
    dt = (time.perf_counter() - t0) / calls
    log(f"drop-in (host arrays, PCIe-inclusive): {dt * 1e3:.2f} ms/call, {sc.n / dt / 1e6:.1f} M particle-updates/s")
    return {"ms_per_call": dt * 1e3, "particle_updates_per_s": sc.n / dt, "calls": calls,
            "note": "stateless sph_hip_update_particles_box on host AoS + transforms, steps 1-6 of the job"}


def batched_leg(nscenes=256, steps=500, tune_steps=5, seed=1024, auto_grid=True):
    scenes = sph_scenes.batched_workload(nscenes, seed)
    descs = [sph_serving.scene_desc(sc, auto_grid=auto_grid) for sc in scenes]
    total = sum(sc.n for sc in scenes)
    log(f"batched: {nscenes} scenes, {total} particles, {steps} steps")
    # The grid tuner (block size x cell scale 1 / 0.5) on the initial states, in
    # a batch of its own: reported, while the timed job keeps cell scale 1.
    with sph_amd.Batch(descs) as g:
        for k, sc in enumerate(scenes):
            g.upload(k, sc.pos, sc.vel)
        grid_initial = g.autotune(tune_steps, grid=True)
    log(f"grid autotune, initial states: {grid_initial}")
    with sph_amd.Batch(descs) as b:
        for k, sc in enumerate(scenes):
            b.upload(k, sc.pos, sc.vel)
        tune = b.autotune(tune_steps)
        log(f"autotune: {tune}")
        b.step_timed(2)  # graph capture
        for k, sc in enumerate(scenes):  # the timed job starts from the initial states
            b.upload(k, sc.pos, sc.vel)
        b.sync()
        t0 = time.perf_counter()
        ms = b.step_timed(steps)
        wall = time.perf_counter() - t0
        grid_final = b.autotune(tune_steps, grid=True)  # the same tuner on the job's end states
    log(f"grid autotune, end of the job: {grid_final}")
    batched_rate = nscenes * steps / wall
    log(f"batched: {wall:.3f} s, {batched_rate:.0f} scene-steps/s, step p50 {np.percentile(ms, 50):.3f} "
        f"p99 {np.percentile(ms, 99):.3f} ms")
    # the same scenes, one after another, each on its own single-scene context
    seq = 0.0
    for sc, d in zip(scenes, descs):
        with sph_amd.Context(sc.n, d.settings, d.box, sc.dt, auto_grid=auto_grid) as ctx:
            ctx.upload(sc.pos, sc.vel)
            ctx.step(1)  # graph capture
            ctx.upload(sc.pos, sc.vel)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step(steps)
            ctx.sync()
            seq += time.perf_counter() - t0
    seq_rate = nscenes * steps / seq
    log(f"sequential single-scene stepping: {seq:.3f} s, {seq_rate:.0f} scene-steps/s")
    return {
        "workload": f"{nscenes} scenes, particles log-uniform in [1k, 64k] (total {total}), h = 2.1 spacing, "
                    f"viscosity U[0.5,2], gas constant U[0.5,3], {steps} steps, one launch per pass, graphed, "
                    f"cell size {'automatic per scene' if auto_grid else 'h'}",
        "scene_steps_per_s": batched_rate,
        "particle_updates_per_s": total * steps / wall,
        "step_ms_p50": float(np.percentile(ms, 50)),
        "step_ms_p99": float(np.percentile(ms, 99)),
        "sequential_scene_steps_per_s": seq_rate,
        "speedup_vs_sequential": batched_rate / seq_rate,
        "autotune": tune,
        "autotune_grid": {"initial_states": grid_initial, "end_of_job": grid_final,
                          "note": "block size x cell scale {1, 0.5} for all scenes at once, reported only; "
                                  "the timed job chooses each scene's cell size from its own state"},
    }


def serving_leg(rates, window, budget, pregen=240, max_queued=64, drain_s=60.0):
    """The serving loop under Poisson streams at several rates: every request
    on a scene of its own drawn size (the first `pregen` requests' scenes are
    generated before the windows, shared by all rates); the server refuses
    arrivals while max_queued requests wait; after each window the loop
    drains, so latencies cover every accepted request."""
    t0 = time.perf_counter()
    scenes = sph_serving.pregenerate(pregen)
    log(f"serving: {len(scenes)} request scenes generated in {time.perf_counter() - t0:.1f} s")
    out = []
    for r in rates:
        res = sph_serving.run_stream(rate=r, duration=window, scenes=scenes, memory_budget=budget,
                                     max_queued=max_queued, drain_s=drain_s)
        log(f"serving at {r} req/s: " + ", ".join(f"{k}={v:.4g}" if isinstance(v, float) else f"{k}={v}"
                                                   for k, v in res.items()))
        out.append(res)
    return {
        "requests": "Poisson arrivals; each request's particles log-uniform [2k, 500k] (its own scene at that "
                    "size), steps U[10, 1000], dam-break or droplet; server queue limit "
                    f"{max_queued} (arrivals beyond are refused and counted), backfill admission, drained "
                    "after the window",
        "memory_budget_bytes": budget,
        "window_s_per_rate": window,
        "rates": out,
    }


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
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--serving-rates", default="50,500,2000", help="request rates (per second) to replay")
    ap.add_argument("--serving-window", type=float, default=5.0, help="seconds of arrivals per rate")
    ap.add_argument("--serving-budget", type=int, default=2 << 30, help="device bytes for admitted scenes")
    ap.add_argument("--no-reset", action="store_true",
                    help="time the steps that follow the warm-up instead of restarting from the initial state "
                         "(to profile late-job steps: --warmup 950 --steps 10)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step's kernels directly instead of replaying a HIP graph (same kernels; "
                         "for kernel tracers that do not follow graph replays)")
    ap.add_argument("--fixed-grid", action="store_true",
                    help="keep cells of h for the whole job instead of the automatic cell size")
    ap.add_argument("--no-job", action="store_true",
                    help="skip the whole 1000-step job that is timed beside a shorter window")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path", file=sys.stderr)
        sys.exit(2)

    sc = sph_scenes.workload(args.workload)
    log(f"workload {sc.name}: n={sc.n} h={sc.h} dt={sc.dt}")
    s = sph_amd.settings(*sc.settings_args())
    b = sph_amd.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    auto = not args.fixed_grid
    ctx = sph_amd.Context(sc.n, s, b, sc.dt, auto_grid=auto)
    ctx.set_graph(not args.no_graph)
    ctx.upload(sc.pos, sc.vel)
    ctx.step(args.warmup)
    ctx.sync()
    if not args.no_reset:
        ctx.upload(sc.pos, sc.vel)  # the timed job starts from the initial state
        ctx.sync()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = sc.n * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step, {value / 1e9:.3f} G particle-updates/s")

    window = "early" if args.steps <= 100 else "job"
    roofline = None
    if not args.no_profile:
        ctx.upload(sc.pos, sc.vel)  # same job again, with events around every kernel
        prof = ctx.profile(args.steps)
        log("per-kernel mean ms (HIP events): " + ", ".join(f"{k}={v:.4f}" for k, v in prof.items()))
        roofline = roofline_of(prof, sc.n, value, window)

    job = None
    if not args.no_job and args.steps != 1000 and not args.no_reset:
        # the representative figure: the whole 1000-step job from the initial
        # state (lattice, collapse, clumps), beside the short window above
        ctx.upload(sc.pos, sc.vel)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(1000)
        ctx.sync()
        jt = time.perf_counter() - t0
        log(f"whole job: 1000 steps in {jt:.3f} s, {jt:.4f} ms/step, {sc.n * 1000 / jt / 1e9:.3f} G particle-updates/s")
        job = {"steps": 1000, "ms_per_step": jt, "value": sc.n * 1000 / jt, "unit": "particle-updates/s"}
        if not args.no_profile:
            ctx.upload(sc.pos, sc.vel)
            jprof = ctx.profile(1000)
            log("whole job per-kernel mean ms: " + ", ".join(f"{k}={v:.4f}" for k, v in jprof.items()))
            job["roofline"] = roofline_of(jprof, sc.n, job["value"], "job")
    ctx.close()

    # measured HBM bandwidth of a plain copy kernel (read + write), beside the 8 TB/s peak
    copy_bytes = args.copy_bytes if args.copy_bytes > 0 else 1 << 30
    ms = sph_amd.copy_bench(copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / ms / 1e6
    log(f"copy kernel: {copy_bytes} B in {ms:.4f} ms = {copy_gbs:.1f} GB/s (read+write)")
    if roofline is not None:
        roofline["copy_kernel_GBs"] = round(copy_gbs, 1)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, multithreaded) ...")
        cpu = cpu_baseline(sc, job_states(sc, args.steps))
        log(f"cpu baseline: {cpu['value'] / 1e6:.3f} M particle-updates/s")

    dropin = dropin_leg(sc)
    batched = None if args.no_batched else batched_leg(steps=args.batched_steps, auto_grid=auto)
    serving = None
    if not args.no_serving:
        rates = [float(x) for x in args.serving_rates.split(",") if x]
        serving = serving_leg(rates, args.serving_window, args.serving_budget)

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
            "cell_size": "automatic (h/2 or h, chosen every 64 steps)" if auto else "h",
        },
        "roofline": roofline,
        "job": job,
        "cpu_baseline": cpu,
        "dropin_pcie_inclusive": dropin,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
