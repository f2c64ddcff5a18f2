"""GPU tests of the serving loop: requests that join and leave a running
batch give results bit-identical to the same scene stepped alone, under a
memory budget that forces queueing."""
import numpy as np
import pytest

import sph_scenes
import sph_serving

pytestmark = pytest.mark.gpu


def _alone(sph, sc, steps):
    with sph.Context(sc.n, sph.settings(*sc.settings_args()),
                     sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt) as ctx:
        ctx.upload(sc.pos, sc.vel)
        ctx.step(steps)
        return ctx.download()


def _same(a, b):
    return all(a[k].tobytes() == b[k].tobytes() for k in ("pos", "vel", "rho", "pressure", "ids"))


def test_server_results_equal_alone_with_arrivals_and_departures(sph):
    scenes = [sph_scenes.request_scene(3000, "dambreak", 1), sph_scenes.request_scene(9000, "droplet", 2),
              sph_scenes.request_scene(20000, "dambreak", 3), sph_scenes.request_scene(5000, "droplet", 4),
              sph_scenes.blobs(nblobs=4, per_blob=1500)]
    steps = [37, 12, 25, 50, 18]
    with sph.Server(max_steps_per_tick=5) as srv:
        ids = [srv.submit(sph_serving.scene_desc(scenes[0]), scenes[0].pos, scenes[0].vel, steps[0]),
               srv.submit(sph_serving.scene_desc(scenes[1]), scenes[1].pos, scenes[1].vel, steps[1])]
        srv.tick()
        srv.tick()
        # joins while the first two run
        ids.append(srv.submit(sph_serving.scene_desc(scenes[2]), scenes[2].pos, scenes[2].vel, steps[2]))
        srv.tick()
        ids.append(srv.submit(sph_serving.scene_desc(scenes[3]), scenes[3].pos, scenes[3].vel, steps[3]))
        ids.append(srv.submit(sph_serving.scene_desc(scenes[4]), scenes[4].pos, scenes[4].vel, steps[4]))
        for _ in range(200):
            if all(srv.status(r)[0] == sph.SPH_REQ_DONE for r in ids):
                break
            srv.tick()
        got = [srv.fetch(r) for r in ids]
        st = srv.stats()
    assert st["completed"] == len(ids)
    for k, sc in enumerate(scenes):
        assert _same(got[k], _alone(sph, sc, steps[k])), f"request {k} ({sc.name}) differs from alone"


def test_server_memory_budget_queues_and_still_matches(sph):
    scenes = [sph_scenes.request_scene(12000, "dambreak", 11 + k) for k in range(4)]
    steps = [9, 14, 6, 11]
    # the charge of one scene (all four have the same count and geometry,
    # hence the same grid), then a budget with room for two of them
    with sph.Server(max_steps_per_tick=1) as srv:
        srv.submit(sph_serving.scene_desc(scenes[0]), scenes[0].pos, scenes[0].vel, 3)
        srv.tick()  # admitted, one of its three steps run
        one = srv.stats()["used_bytes"]
    assert one > 116 * 12000
    budget = int(2.5 * one)
    with sph.Server(memory_budget=budget, max_steps_per_tick=4) as srv:
        ids = [srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, s) for sc, s in zip(scenes, steps)]
        srv.tick()
        states = [srv.status(r)[0] for r in ids]
        assert states.count(sph.SPH_REQ_QUEUED) >= 1, states
        assert srv.stats()["used_bytes"] <= budget
        for _ in range(100):
            if all(srv.status(r)[0] == sph.SPH_REQ_DONE for r in ids):
                break
            srv.tick()
            assert srv.stats()["used_bytes"] <= budget
        got = [srv.fetch(r) for r in ids]
    for k, sc in enumerate(scenes):
        assert _same(got[k], _alone(sph, sc, steps[k])), f"request {k} differs from alone"


def test_server_rejects_what_never_fits_and_bad_calls(sph):
    sc = sph_scenes.request_scene(4000, "droplet", 5)
    with sph.Server(memory_budget=1 << 16) as srv:
        with pytest.raises(sph.SphError) as e:
            srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, 5)
        assert e.value.code == sph.SPH_E_LIMIT
    with sph.Server() as srv:
        rid = srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, 3)
        with pytest.raises(sph.SphError) as e:
            srv.fetch(rid)  # not finished
        assert e.value.code == sph.SPH_E_STATE
        with pytest.raises(sph.SphError):
            srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, 0)


def test_server_stream_smoke(sph):
    pool = [sph_scenes.request_scene(n, k, 7 + i) for i, (n, k) in
            enumerate([(2500, "dambreak"), (6000, "droplet"), (15000, "dambreak"), (4000, "droplet")])]
    res = sph_serving.run_stream(rate=200.0, duration=1.5, pool=pool, seed=3, max_queue=16)
    assert res["completed"] >= 1
    assert res["batch_steps"] > 0 and res["step_ms_p99"] >= res["step_ms_p50"] > 0
    assert res["particle_updates_per_s"] > 0


@pytest.mark.parametrize("stage", ["relayout", "upload"])
def test_server_failed_admission_requeues_and_keeps_running_scenes(sph, stage):
    """A tick whose admission fails (injected at the relayout's last fallible
    point, or at the upload of the first joining scene) returns the error,
    puts the joining requests back at the head of the queue with their
    memory charge released, and leaves the running scene untouched; the next
    tick admits them and every result still equals the scene stepped alone.
    The first admission grows the batch from nothing, so the relayout case
    covers the grow path's roll-back; the second admission joins a running
    scene."""
    code = sph.SPH_SERVER_FAIL_RELAYOUT if stage == "relayout" else sph.SPH_SERVER_FAIL_UPLOAD
    scenes = [sph_scenes.request_scene(4000, "dambreak", 21), sph_scenes.request_scene(7000, "droplet", 22),
              sph_scenes.request_scene(3000, "dambreak", 23)]
    steps = [12, 9, 7]
    with sph.Server(max_steps_per_tick=3) as srv:
        # first admission: two joining scenes on an empty batch
        a = srv.submit(sph_serving.scene_desc(scenes[0]), scenes[0].pos, scenes[0].vel, steps[0])
        b = srv.submit(sph_serving.scene_desc(scenes[1]), scenes[1].pos, scenes[1].vel, steps[1])
        srv.inject_failure(code)
        with pytest.raises(sph.SphError):
            srv.tick()
        st = srv.stats()
        assert st["queued"] == 2 and st["active"] == 0 and st["used_bytes"] == 0, st
        assert [srv.status(r)[0] for r in (a, b)] == [sph.SPH_REQ_QUEUED] * 2
        srv.tick()
        assert [srv.status(r)[:2] for r in (a, b)] == [(sph.SPH_REQ_ACTIVE, 3)] * 2
        charged = srv.stats()["used_bytes"]
        # second admission: joins the two running scenes
        c = srv.submit(sph_serving.scene_desc(scenes[2]), scenes[2].pos, scenes[2].vel, steps[2])
        srv.inject_failure(code)
        with pytest.raises(sph.SphError):
            srv.tick()
        st = srv.stats()
        assert st["queued"] == 1 and st["active"] == 2 and st["used_bytes"] == charged, st
        assert srv.status(c)[0] == sph.SPH_REQ_QUEUED
        assert [srv.status(r)[:2] for r in (a, b)] == [(sph.SPH_REQ_ACTIVE, 3)] * 2
        ids = [a, b, c]
        for _ in range(50):
            if all(srv.status(r)[0] == sph.SPH_REQ_DONE for r in ids):
                break
            srv.tick()
        got = [srv.fetch(r) for r in ids]
        assert srv.stats()["used_bytes"] == 0
    for k, sc in enumerate(scenes):
        assert _same(got[k], _alone(sph, sc, steps[k])), f"request {k} ({sc.name}) differs from alone"
