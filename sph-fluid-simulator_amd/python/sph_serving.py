"""Replays a Poisson stream of scene-advance requests against the serving
loop (sph_amd.Server) in real time and reports latency and throughput.

Requests follow sph_scenes.request_stream: particle counts log-uniform in
[2k, 500k], steps uniform in [10, 1000], dam-break or droplet geometry.
Every request gets a scene of its own drawn size, geometry and seed.  The
i-th request of a stream is the same at every rate (the stream draws its
fields in the same order), so the scenes of the first requests are
generated once (pregenerate), before any window, and shared by all rates; a later
request's scene is generated when it arrives and the server has room for
it.  The server bounds its queue (max_queued): a request arriving while the
queue is full is refused by the server and counted as rejected.  After the
arrival window the loop keeps ticking until every accepted request has
finished (or `drain_s` runs out), so latencies cover every accepted
request.
"""
from __future__ import annotations

import time

import numpy as np

import sph_amd
import sph_scenes


def pregenerate(count: int, seed: int = 1024, rate: float = 1000.0, **sizes):
    """Scenes of the first `count` requests of the stream drawn with `seed`
    (the same at every rate; `sizes`: request_stream's count and step
    ranges)."""
    stream = sph_scenes.request_stream(rate, 1e9, seed, max_requests=count, **sizes)
    return [sph_scenes.request_scene(n, kind, sd) for _, n, _, kind, sd in stream]


def scene_desc(sc, auto_grid=False):
    return sph_amd.scene_desc(sc.n, sph_amd.settings(*sc.settings_args()),
                              sph_amd.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt,
                              auto_grid=auto_grid)


def run_stream(rate: float, duration: float, scenes=(), seed: int = 1024, memory_budget: int = 0,
               max_scenes: int = 0, max_steps_per_tick: int = 8, max_queued: int = 64, graph: bool = True,
               drain_s: float = 60.0, backfill: bool = True, **sizes):
    """Feeds the arrivals of a Poisson stream at `rate` per second for
    `duration` seconds of wall time (scenes[i], when given, is request i's
    scene, from pregenerate with the same seed and sizes), drains, and
    returns the measurements."""
    stream = sph_scenes.request_stream(rate, duration, seed, **sizes)
    with sph_amd.Server(memory_budget, max_scenes, max_steps_per_tick, graph, max_queued, backfill) as srv:
        pending = {}
        lat, wait, counts = [], [], []
        submitted = rejected = 0
        i = 0
        t_start = time.perf_counter()
        t_window = None
        while True:
            now = time.perf_counter() - t_start
            if now < duration:
                queued = sum(1 for st in pending.values() if st == sph_amd.SPH_REQ_QUEUED)
                while i < len(stream) and stream[i][0] <= now:
                    _, n, steps, kind, sd = stream[i]
                    if queued >= max_queued > 0:
                        rejected += 1  # the server refuses it: no scene is built for it
                    else:
                        sc = scenes[i] if i < len(scenes) else sph_scenes.request_scene(n, kind, sd)
                        try:
                            rid = srv.submit(scene_desc(sc), sc.pos, sc.vel, steps)
                            pending[rid] = sph_amd.SPH_REQ_QUEUED
                            counts.append(n)
                            submitted += 1
                            queued += 1
                        except sph_amd.SphError as e:  # refused: counted by the server
                            if e.code != sph_amd.SPH_E_STATE:
                                raise
                    i += 1
            else:
                if t_window is None:
                    t_window = now
                if not pending or now >= duration + drain_s:
                    break
            ran = srv.tick()
            for rid in list(pending):
                st, _, t = srv.status(rid)
                pending[rid] = st
                if st == sph_amd.SPH_REQ_DONE:
                    srv.discard(rid)
                    lat.append(t[2] - t[0])
                    wait.append(t[1] - t[0])
                    del pending[rid]
            if ran == 0 and now < duration and i < len(stream):
                time.sleep(max(0.0, min(stream[i][0] - (time.perf_counter() - t_start), 0.002)))
        wall = time.perf_counter() - t_start
        st = srv.stats()
        steps_ms = srv.step_times()

    def pct(v, q):
        return float(np.percentile(v, q)) if len(v) else None

    return {
        "rate_req_s": rate,
        "window_s": round(t_window if t_window is not None else wall, 3),
        "drain_s": round(wall - (t_window if t_window is not None else wall), 3),
        "arrivals": i,
        "submitted": submitted,
        "rejected_queue_full": rejected + int(st["rejected"]),
        "completed": len(lat),
        "unfinished": len(pending),
        "backfilled": int(st["backfilled"]),
        "particles_per_request_mean": float(np.mean(counts)) if counts else None,
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
