"""Replays a Poisson stream of scene-advance requests against the serving
loop (sph_amd.Server) in real time and reports latency and throughput.

Requests follow sph_scenes.request_stream: particle counts log-uniform in
[2k, 500k], steps uniform in [10, 1000], dam-break or droplet geometry.  The
initial states come from a pool of pre-generated request scenes (request k
uses pool entry k mod pool_size with its own step count), so scene
generation stays off the clock.  A client-side bound on queued requests
stands in for admission back-pressure: arrivals beyond it are counted as
dropped.
"""
from __future__ import annotations

import time

import numpy as np

import sph_amd
import sph_scenes


def make_pool(size: int, seed: int = 1024):
    """Request scenes drawn like the stream's requests (count log-uniform in
    [2k, 500k], geometries alternating), pre-generated."""
    rng = np.random.default_rng(seed + 17)
    pool = []
    for k in range(size):
        n = int(round(np.exp(rng.uniform(np.log(2000), np.log(500000)))))
        kind = "dambreak" if k % 2 == 0 else "droplet"
        pool.append(sph_scenes.request_scene(n, kind, seed + 31 * k))
    return pool


def scene_desc(sc, auto_grid=False):
    return sph_amd.scene_desc(sc.n, sph_amd.settings(*sc.settings_args()),
                              sph_amd.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt,
                              auto_grid=auto_grid)


def run_stream(rate: float, duration: float, pool, seed: int = 1024, memory_budget: int = 0,
               max_scenes: int = 0, max_steps_per_tick: int = 8, max_queue: int = 64, graph: bool = True):
    """Feeds arrivals of a Poisson stream at `rate` per second for `duration`
    seconds of wall time and returns the measurements."""
    stream = sph_scenes.request_stream(rate, duration, seed)
    descs = [scene_desc(sc) for sc in pool]
    with sph_amd.Server(memory_budget, max_scenes, max_steps_per_tick, graph) as srv:
        pending = {}
        lat, wait = [], []
        submitted = dropped = 0
        i = 0
        t_start = time.perf_counter()
        while True:
            now = time.perf_counter() - t_start
            if now >= duration:
                break
            queued = sum(1 for st in pending.values() if st == sph_amd.SPH_REQ_QUEUED)
            while i < len(stream) and stream[i][0] <= now:
                _, _, steps, _, _ = stream[i]
                if queued < max_queue:
                    sc = pool[i % len(pool)]
                    rid = srv.submit(descs[i % len(pool)], sc.pos, sc.vel, steps)
                    pending[rid] = sph_amd.SPH_REQ_QUEUED
                    submitted += 1
                    queued += 1
                else:
                    dropped += 1
                i += 1
            ran = srv.tick()
            for rid in list(pending):
                st, _, t = srv.status(rid)
                pending[rid] = st
                if st == sph_amd.SPH_REQ_DONE:
                    srv.discard(rid)
                    lat.append(t[2] - t[0])
                    wait.append(t[1] - t[0])
                    del pending[rid]
            if ran == 0 and i < len(stream):
                time.sleep(max(0.0, min(stream[i][0] - (time.perf_counter() - t_start), 0.002)))
        wall = time.perf_counter() - t_start
        st = srv.stats()
        steps_ms = srv.step_times()

    def pct(v, q):
        return float(np.percentile(v, q)) if len(v) else None

    return {
        "rate_req_s": rate,
        "window_s": round(wall, 3),
        "arrivals": i,
        "submitted": submitted,
        "dropped_queue_full": dropped,
        "completed": len(lat),
        "batch_steps": int(st["steps"]),
        "step_ms_p50": pct(steps_ms, 50),
        "step_ms_p99": pct(steps_ms, 99),
        "particle_updates_per_s": st["particle_steps"] / wall,
        "scene_steps_per_s": st["scene_steps"] / wall,
        "gpu_busy_frac": st["busy_ms"] / (1e3 * wall),
        "request_latency_s_p50": pct(lat, 50),
        "request_latency_s_p99": pct(lat, 99),
        "queue_wait_s_p50": pct(wait, 50),
        "queue_wait_s_p99": pct(wait, 99),
        "memory_budget_bytes": int(st["budget_bytes"]),
    }
