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
    """A short stream of small requests, each on a scene of its own drawn
    size: every accepted request finishes in the drain, and the server's
    queue limit refuses the rest."""
    sizes = {"n_lo": 2000, "n_hi": 15000, "steps_lo": 5, "steps_hi": 40}
    scenes = sph_serving.pregenerate(40, seed=3, **sizes)
    res = sph_serving.run_stream(rate=200.0, duration=1.0, scenes=scenes, seed=3, max_queued=8, **sizes)
    assert res["completed"] == res["submitted"] >= 1 and res["unfinished"] == 0
    assert res["arrivals"] == res["submitted"] + res["rejected_queue_full"]
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


def test_served_requests_match_the_oracle(sph, orc32):
    """Serving workload (dam-break and droplet requests of 3k-100k
    particles, plus one of 400k): every served result equals the same scene
    stepped alone, and that scene's last step matches the oracle fed the
    state one step earlier -- density and pressure within 1e-5, force within
    the scaled bar, and position and velocity bit-exact when the oracle
    integrates with the GPU's force and density (src/sph.cpp:195-275).  The
    400k request, served beside the others, keeps ids, finite in-box
    positions and sampled densities within 1e-5."""
    import parity_util as pu
    reqs = [(3000, "dambreak", 31, 6), (12000, "droplet", 32, 9), (100000, "dambreak", 33, 5),
            (30000, "droplet", 34, 7), (400000, "droplet", 35, 4)]
    scenes = [sph_scenes.request_scene(n, kind, seed) for n, kind, seed, _ in reqs]
    steps = [r[3] for r in reqs]
    with sph.Server(max_steps_per_tick=3) as srv:
        ids = []
        for k, sc in enumerate(scenes):
            ids.append(srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, steps[k]))
            srv.tick()  # arrivals while others run
        for _ in range(100):
            if all(srv.status(r)[0] == sph.SPH_REQ_DONE for r in ids):
                break
            srv.tick()
        got = [srv.fetch(r) for r in ids]
    for k, sc in enumerate(scenes[:4]):
        label = f"request {k} ({sc.name})"
        print(label, flush=True)
        with sph.Context(sc.n, sph.settings(*sc.settings_args()),
                         sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt, keep_force=True) as ctx:
            ctx.upload(sc.pos, sc.vel)
            ctx.step(steps[k] - 1)
            before = ctx.download()
            ctx.step(1)
            p, _ = ctx.download_particles(transforms=False)
            after = ctx.download()
        assert _same(got[k], after), f"{label}: served result differs from the scene stepped alone"
        st = {}
        for key in ("pos", "vel"):
            st[key] = np.zeros_like(before[key])
            st[key][before["ids"].astype(np.int64)] = before[key]
        ai = after["ids"].astype(np.int64)
        nx = {key: np.zeros_like(after[key]) for key in ("pos", "vel", "rho", "pressure")}
        for key in nx:
            nx[key][ai] = after[key]
        f = np.zeros((sc.n, 3), dtype=np.float32)
        f[ai] = p["force"][:, :3]
        s_o = orc32.settings(*sc.settings_args())
        box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
        ref, _, rows, _, _ = pu.oracle_step(orc32, st["pos"], st["vel"], s_o, sc.dt, box_o)
        pu.assert_rho_p(nx["rho"], nx["pressure"], ref["density"], ref["pressure"], s_o, label)
        scale = pu.force_scale(st["pos"], st["vel"], ref["density"], ref["pressure"], rows, s_o)
        pu.assert_force(f, ref["force"][:, :3].astype(np.float64), scale, label)
        q = orc32.particles(st["pos"], st["vel"])
        q["force"] = f
        q["density"] = nx["rho"]
        orc32.integrate(q, s_o, sc.dt, box_o, transforms=False)
        assert np.array_equal(q["position"], nx["pos"]), f"{label}: integrate"
        assert np.array_equal(q["velocity"], nx["vel"]), f"{label}: integrate velocity"
    big, g = scenes[4], got[4]
    ids = g["ids"].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(big.n))
    assert np.isfinite(g["pos"]).all() and np.isfinite(g["vel"]).all()
    lo = np.array(big.box_lo) + big.h
    hi = np.array(big.box_hi) - big.h
    assert np.all(g["pos"] >= lo - 1e-3) and np.all(g["pos"] <= hi + 1e-3)
    s_o = orc32.settings(*big.settings_args())
    assert np.all(g["rho"] >= np.float32(s_o.selfDens))
    with sph.Context(big.n, sph.settings(*big.settings_args()),
                     sph.box(big.box_lo, big.box_hi, big.restitution, big.offset), big.dt) as ctx:
        ctx.upload(big.pos, big.vel)
        ctx.step(steps[4] - 1)
        before = ctx.download()
        ctx.step(1)
        after = ctx.download()
    assert _same(g, after), "400k request: served result differs from the scene stepped alone"
    st = {}
    for key in ("pos", "vel"):
        st[key] = np.zeros_like(before[key])
        st[key][before["ids"].astype(np.int64)] = before[key]
    rho = np.zeros(big.n, dtype=np.float32)
    rho[ids] = g["rho"]
    p = orc32.particles(st["pos"], st["vel"], ids=np.arange(big.n))
    table = orc32.prepare(p, s_o)
    where = np.zeros(big.n, dtype=np.int64)
    where[p["acceleration"][:, 0].astype(np.int64)] = np.arange(big.n)
    sample = np.random.default_rng(3).choice(big.n, size=4000, replace=False)
    rho_ref, _ = orc32.density_subset(p, table, s_o, where[sample])
    orc32.free_table(table)
    rel = np.abs(rho[sample].astype(np.float64) - rho_ref) / rho_ref
    assert rel.max() <= 1e-5, rel.max()


def _charge(sph, sc):
    """Device bytes the server charges for scene sc."""
    with sph.Server(max_steps_per_tick=1) as srv:
        srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, 2)
        srv.tick()
        return srv.stats()["used_bytes"]


@pytest.mark.parametrize("backfill", [True, False])
def test_server_backfill_never_delays_the_head(sph, backfill):
    """A long request runs; the next (large) one does not fit beside it, a
    short small one behind it does: with backfill the small one starts at
    once (it ends before the large one could start), without it waits its
    turn.  The large one starts at the same batch step either way, and every
    result equals its scene stepped alone."""
    a = sph_scenes.request_scene(20000, "dambreak", 41)
    bg = sph_scenes.request_scene(20000, "droplet", 42)
    c = sph_scenes.request_scene(3000, "droplet", 43)
    ca, cb, cc = _charge(sph, a), _charge(sph, bg), _charge(sph, c)
    budget = ca + max(cb, cc) - 1 if cc < cb else ca + cc
    assert ca + cc <= budget < ca + cb, "premise: c fits beside a, b does not"
    steps = [24, 6, 6]
    with sph.Server(memory_budget=budget, max_steps_per_tick=3, backfill=backfill) as srv:
        ids = [srv.submit(sph_serving.scene_desc(a), a.pos, a.vel, steps[0])]
        srv.tick()
        ids.append(srv.submit(sph_serving.scene_desc(bg), bg.pos, bg.vel, steps[1]))
        ids.append(srv.submit(sph_serving.scene_desc(c), c.pos, c.vel, steps[2]))
        srv.tick()
        states = [srv.status(r)[0] for r in ids]
        assert states[1] == sph.SPH_REQ_QUEUED
        assert states[2] == (sph.SPH_REQ_ACTIVE if backfill else sph.SPH_REQ_QUEUED), states
        assert srv.stats()["backfilled"] == (1 if backfill else 0)
        for _ in range(100):
            if all(srv.status(r)[0] == sph.SPH_REQ_DONE for r in ids):
                break
            srv.tick()
            assert srv.stats()["used_bytes"] <= budget
        admit_b = srv.status(ids[1])[2][1]
        done_a = srv.status(ids[0])[2][2]
        got = [srv.fetch(r) for r in ids]
    assert admit_b >= done_a  # b starts once a has left, in both modes
    for k, sc in enumerate((a, bg, c)):
        assert _same(got[k], _alone(sph, sc, steps[k])), f"request {k} differs from alone"


def test_server_queue_limit_rejects_and_counts(sph):
    sc = sph_scenes.request_scene(2500, "droplet", 51)
    with sph.Server(max_queued=2) as srv:
        srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, 3)
        srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, 3)
        with pytest.raises(sph.SphError) as e:
            srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, 3)
        assert e.value.code == sph.SPH_E_STATE
        st = srv.stats()
        assert st["rejected"] == 1 and st["queued"] == 2 and st["submitted"] == 2
        srv.tick()  # both admitted: the queue has room again
        srv.submit(sph_serving.scene_desc(sc), sc.pos, sc.vel, 3)
        assert srv.stats()["rejected"] == 1
