"""GPU tests of the multi-scene batch: every scene of a batch steps
bit-identically to the same scene stepped alone, batches keep parity with
the oracle, and dense cells (sorted in-cell by A index) stay deterministic."""
import numpy as np
import pytest

import parity_util as pu
import sph_scenes

pytestmark = pytest.mark.gpu
F32 = np.float32


def _desc(sph, sc):
    return sph.scene_desc(sc.n, sph.settings(*sc.settings_args()),
                          sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt)


def _mixed_scenes():
    scenes = sph_scenes.batched_workload(nscenes=5, seed=77)
    app = sph_scenes.reference_app_scene(9)  # reference CPU box, no ceiling
    scenes.append(app)
    w0 = sph_scenes.workload(0)  # open ceiling, different h and dt
    scenes.append(w0)
    # clumps: its long walks split into parts (queued) while the other
    # scenes' short tasks run in the waves' static share
    scenes.append(sph_scenes.blobs(nblobs=4, per_blob=3000))
    return scenes


def _state(d):
    return (d["pos"].tobytes() + d["vel"].tobytes() + d["rho"].tobytes() + d["pressure"].tobytes()
            + d["ids"].tobytes())


def test_batch_scene_equals_scene_alone(sph):
    scenes = _mixed_scenes()
    steps = 60
    with sph.Batch([_desc(sph, sc) for sc in scenes]) as b:
        for k, sc in enumerate(scenes):
            b.upload(k, sc.pos, sc.vel)
        b.step(steps)
        mixed = [_state(b.download(k)) for k in range(len(scenes))]
    for k, sc in enumerate(scenes):
        with sph.Context(sc.n, sph.settings(*sc.settings_args()),
                         sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt) as ctx:
            ctx.upload(sc.pos, sc.vel)
            ctx.step(steps)
            alone = _state(ctx.download())
        assert alone == mixed[k], f"scene {k} ({sc.name}) differs alone vs batched"
    # premise: the clumped scene has split tasks (only a split task makes the
    # union and per-lane walks differ in the last bits)
    sc = scenes[-1]
    walks = []
    for lane_walk in (False, True):
        with sph.Context(sc.n, sph.settings(*sc.settings_args()),
                         sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt,
                         lane_walk=lane_walk) as ctx:
            ctx.upload(sc.pos, sc.vel)
            ctx.step(1)
            walks.append(_state(ctx.download()))
    assert walks[0] != walks[1], "premise: the blob scene splits some wave tasks"


def test_batch_order_of_scenes_does_not_matter(sph):
    scenes = _mixed_scenes()[:4]
    order = [2, 0, 3, 1]
    res = {}
    for perm in (list(range(4)), order):
        with sph.Batch([_desc(sph, scenes[k]) for k in perm]) as b:
            for slot, k in enumerate(perm):
                b.upload(slot, scenes[k].pos, scenes[k].vel)
            b.step(25)
            for slot, k in enumerate(perm):
                res.setdefault(k, []).append(_state(b.download(slot)))
    for k in range(4):
        assert res[k][0] == res[k][1]


def test_batch_teacher_forced_parity(sph, orc32):
    """Three batched scenes after 10 steps: the next batched step against
    the oracle fed the same state -- density, pressure and force."""
    scenes = sph_scenes.batched_workload(nscenes=3, seed=5)
    descs = [sph.scene_desc(sc.n, sph.settings(*sc.settings_args()),
                            sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt, keep_force=True)
             for sc in scenes]
    with sph.Batch(descs) as b:
        for k, sc in enumerate(scenes):
            b.upload(k, sc.pos, sc.vel)
        b.step(10)
        before = [b.download(k) for k in range(3)]
        b.step(1)
        after = [b.download(k) for k in range(3)]
        forces = [b.download_particles(k, transforms=False)[0]["force"] for k in range(3)]
    for k, sc in enumerate(scenes):
        st = {key: np.zeros_like(v) for key, v in before[k].items() if key != "ids"}
        for key in st:
            st[key][before[k]["ids"].astype(np.int64)] = before[k][key]
        nx = {key: np.zeros_like(v) for key, v in after[k].items() if key != "ids"}
        for key in nx:
            nx[key][after[k]["ids"].astype(np.int64)] = after[k][key]
        s_o = orc32.settings(*sc.settings_args())
        box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
        ref, _, rows, _, _ = pu.oracle_step(orc32, st["pos"], st["vel"], s_o, sc.dt, box_o)
        pu.assert_rho_p(nx["rho"], nx["pressure"], ref["density"], ref["pressure"], s_o, sc.name)
        f = np.zeros((sc.n, 3), dtype=F32)
        f[after[k]["ids"].astype(np.int64)] = forces[k]
        scale = pu.force_scale(st["pos"], st["vel"], ref["density"], ref["pressure"], rows, s_o)
        pu.assert_force(f, ref["force"][:, :3].astype(np.float64), scale, sc.name)


def _dense_cluster(n, h, seed):
    """n particles inside one h-sized cube away from the axes: one grid cell
    far above the counting-rank limit, so its order comes from the in-cell
    sort."""
    rng = np.random.default_rng(seed)
    return (rng.uniform(0.0, 0.9 * h, size=(n, 3)) + np.array([0.71, 0.53, 0.37])).astype(F32)


def test_dense_cell_parity_and_determinism(sph, orc32):
    h = 0.05
    pos = _dense_cluster(3000, h, 11)
    args = (7.4e-4, 1000.0, 1.0, 1.04, h, -9.8, 0.2)
    # neighbour multisets bit-exact and densities within tolerance
    s_g = sph.settings(*args)
    s_o = orc32.settings(*args)
    goff, gidx = sph.neighbors(sph.make_particles(pos), s_g)
    ref, _, rows, _, _ = pu.oracle_step(orc32, pos, np.zeros_like(pos), s_o, 0.001,
                                        orc32.box((-2, -2, -2), (2, 2, 2)))
    for i in range(len(pos)):
        assert np.array_equal(gidx[goff[i]:goff[i + 1]].astype(np.int64), rows[i])
    inp = sph.make_particles(pos)
    inp["acceleration"][:, 0] = np.arange(len(pos), dtype=F32)
    out, _ = sph.update_particles(inp, s_g, 0.001, sph.box((-2, -2, -2), (2, 2, 2)))
    g, _ = pu.by_id(out)
    pu.assert_rho_p(g["density"], g["pressure"], ref["density"], ref["pressure"], s_o, "dense")
    # resident runs through the big-cell sort are repeatable bit for bit
    states = []
    for _ in range(2):
        with sph.Context(len(pos), s_g, sph.box((-2, -2, -2), (2, 2, 2)), 0.001) as ctx:
            ctx.upload(pos)
            ctx.step(8)
            states.append(_state(ctx.download()))
    assert states[0] == states[1]


def test_batched_workload_runs(sph):
    scenes = sph_scenes.batched_workload(nscenes=32, seed=3)
    with sph.Batch([_desc(sph, sc) for sc in scenes]) as b:
        for k, sc in enumerate(scenes):
            b.upload(k, sc.pos, sc.vel)
        b.step(20)
        b.sync()
        for k, sc in enumerate(scenes):
            d = b.download(k)
            assert np.array_equal(np.sort(d["ids"]), np.arange(sc.n))
            assert np.isfinite(d["pos"]).all()


def test_autotune_restores_state_and_keeps_bits(sph):
    scenes = sph_scenes.batched_workload(nscenes=6, seed=21) + [sph_scenes.blobs(nblobs=3, per_blob=1200)]
    res = []
    for tune in (True, False):
        with sph.Batch([_desc(sph, sc) for sc in scenes]) as b:
            for k, sc in enumerate(scenes):
                b.upload(k, sc.pos, sc.vel)
            if tune:
                t = b.autotune(3)
                assert t["stencil_block"] in (64, 128, 256) and t["lane_walk"] == 0
                assert t["trials"] == 3 and min(t["trial_ms"][:3]) == t["ms_per_step"] > 0
            ms = b.step_timed(15)
            assert ms.shape == (15,) and np.all(ms > 0)
            res.append([_state(b.download(k)) for k in range(len(scenes))])
    assert res[0] == res[1]
