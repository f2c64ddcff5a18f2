"""GPU tests of grids finer than h: cell_scale in [1/3, 1) gives cells of
edge h (1 + 2^-10) cell_scale and a stencil of R = ceil(1 / cell_scale)
cells each way.  One step matches the oracle (density, pressure, force,
also with slow lanes whose bucket multiplicity is 2), the union and
per-lane walks stay bit-identical, the 1M-particle scene keeps its sampled
densities, and a fine-grid scene steps bit-identically alone and in a batch
beside scenes on other grids."""
import numpy as np
import pytest

import parity_util as pu
import sph_scenes

pytestmark = pytest.mark.gpu
F32 = np.float32
SCALES = [0.5, 1.0 / 3.0]


def _ctx(sph, sc, cell_scale, keep_force=False, lane_walk=False):
    ctx = sph.Context(sc.n, sph.settings(*sc.settings_args()),
                      sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt,
                      cell_scale=cell_scale, keep_force=keep_force, lane_walk=lane_walk)
    ctx.upload(sc.pos, sc.vel)
    return ctx


def _by_id(d):
    ids = d["ids"].astype(np.int64)
    out = {}
    for k in ("pos", "vel", "rho", "pressure"):
        a = np.zeros_like(d[k])
        a[ids] = d[k]
        out[k] = a
    return out


def _bytes(d):
    return b"".join(d[k].tobytes() for k in ("pos", "vel", "rho", "pressure", "ids"))


def _one_step_parity(sph, orc32, sc, cell_scale, label):
    """One GPU step from the scene's state against the oracle's step:
    density and pressure within 1e-5 relative, force within the scaled
    tolerance."""
    s_o = orc32.settings(*sc.settings_args())
    box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    ref, _, rows, _, _ = pu.oracle_step(orc32, sc.pos, sc.vel, s_o, sc.dt, box_o)
    with _ctx(sph, sc, cell_scale, keep_force=True) as ctx:
        ctx.step(1)
        p, _ = ctx.download_particles(transforms=False)
        d = ctx.download()
    ids = d["ids"].astype(np.int64)
    nxt = _by_id(d)
    f = np.zeros((sc.n, 3), dtype=F32)
    f[ids] = p["force"][:, :3]
    pu.assert_rho_p(nxt["rho"], nxt["pressure"], ref["density"], ref["pressure"], s_o, label)
    scale = pu.force_scale(sc.pos, sc.vel, ref["density"], ref["pressure"], rows, s_o)
    pu.assert_force(f, ref["force"][:, :3].astype(np.float64), scale, label)


@pytest.mark.parametrize("cell_scale", SCALES)
def test_fine_grid_teacher_forced_steps(sph, orc32, cell_scale):
    """Workload 0 on a fine grid: after 0 and 25 GPU steps, the next GPU
    step matches the oracle's step from the same state."""
    sc = sph_scenes.workload(0)
    s_o = orc32.settings(*sc.settings_args())
    box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    with _ctx(sph, sc, cell_scale) as ctx:
        done = 0
        for k in (0, 25):
            ctx.step(k - done)
            st = _by_id(ctx.download())
            ctx.step(1)
            done = k + 1
            nxt = _by_id(ctx.download())
            ref, _, _, _, _ = pu.oracle_step(orc32, st["pos"], st["vel"], s_o, sc.dt, box_o)
            pu.assert_rho_p(nxt["rho"], nxt["pressure"], ref["density"], ref["pressure"], s_o,
                            f"scale {cell_scale:.3f} step {k}")


@pytest.mark.parametrize("cell_scale", SCALES)
def test_fine_grid_force_and_slow_lanes(sph, orc32, cell_scale):
    """Force parity on workload 0 and on a clump whose slow lanes count
    neighbours twice (m_ij = 2)."""
    _one_step_parity(sph, orc32, sph_scenes.workload(0), cell_scale, f"w0 scale {cell_scale:.3f}")
    sc = pu.collision_clump()
    _one_step_parity(sph, orc32, sc, cell_scale, f"clump scale {cell_scale:.3f}")


def _dense_slab(cell_scale, h=0.05):
    """A slab 32 x 2 x 1 grid cells (cell edge h (1 + 2^-10) cell_scale)
    holding 2 x 2 x 2 particles per cell: in cell order a wave's 64 slots
    span 8 cells in x, so its union of stencils is twice its own and it
    takes the union walk; its few stencil rows keep every walk below the
    part size."""
    H = h * (1.0 + 2.0 ** -10) * cell_scale
    a = 0.5 * H
    pos = sph_scenes.lattice((64, 4, 2), a, [-0.5 + 0.25 * a, 0.1 + 0.25 * a, 0.2 + 0.25 * a], 0.01 * a, 17)
    mass = float(F32(0.02 * (h / 0.15) ** 3))
    # a tiny time step: the slab is far denser than its rest density and
    # would fly apart, changing which tasks split, within a few steps
    # a box tight around it: a scene's grid holds at most 64 cells per
    # particle, and a wider box would coarsen these cells
    return sph_scenes.Scene("dense-slab", pos, np.zeros_like(pos), mass, 1000.0, 1.0, 1.04, h, -9.8, 0.2,
                            1e-6, (-0.6, 0.0, 0.1), (0.2, 0.25, 0.35), steps=10)


@pytest.mark.parametrize("cell_scale", SCALES)
def test_fine_grid_union_matches_lane_walk(sph, cell_scale):
    """The union and per-lane walks give bit-identical states step after
    step on a fine grid: on the 8k cube and on a dense slab whose waves span
    8 cells in x (premise: union walks occur there and no task is split),
    and on the slow-lane clump (its waves all hold a slow lane, so both runs
    walk per lane; on fine grids its tasks are split into parts)."""
    scenes = ((sph_scenes.workload(0), 15), (pu.collision_clump(), 2), (_dense_slab(cell_scale), 2))
    unions = 0
    for sc, every in (scenes[0], scenes[2]):
        for lane_walk in (False, True):
            with _ctx(sph, sc, cell_scale, lane_walk=lane_walk) as ctx:
                for _ in range(3):  # every compared state, both runs
                    dump = ctx.walk_dump()
                    unions += 0 if lane_walk else int(dump["union"].sum())
                    assert not dump["split"].any(), f"{sc.name}: premise, no task is split"
                    ctx.step(every)
    assert unions > 0, "premise: some waves take the union walk"
    for sc, every in scenes:
        states = []
        for lane_walk in (False, True):
            with _ctx(sph, sc, cell_scale, lane_walk=lane_walk) as ctx:
                out = []
                for _ in range(3):
                    ctx.step(every)
                    out.append(_bytes(ctx.download()))
            states.append(out)
        for k, (u, l) in enumerate(zip(*states)):
            assert u == l, f"{sc.name}: walks differ after {every * (k + 1)} steps"


def test_fine_grid_split_tasks(sph, orc32):
    """Clumped scene with split wave tasks on a fine grid: repeated runs are
    bit-identical and the densities match the oracle."""
    sc = sph_scenes.blobs()
    runs = []
    for _ in range(2):
        with _ctx(sph, sc, 0.5) as ctx:
            ctx.step(1)
            d1 = ctx.download()
            ctx.step(4)
            runs.append(_bytes(d1) + _bytes(ctx.download()))
    assert runs[0] == runs[1]
    s_o = orc32.settings(*sc.settings_args())
    ref, _, _, _, _ = pu.oracle_step(orc32, sc.pos, sc.vel, s_o, sc.dt,
                                     orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset))
    u = _by_id(d1)
    pu.assert_rho_p(u["rho"], u["pressure"], ref["density"], ref["pressure"], s_o, "blobs scale 0.5")


def test_fine_grid_workload2_sampled(sph, orc32):
    """1,000,000 particles on the half-h grid: densities of 4000 sampled
    particles after 4 steps against the oracle; ids stay a permutation."""
    sc = sph_scenes.workload(2)
    s_o = orc32.settings(*sc.settings_args())
    with _ctx(sph, sc, 0.5) as ctx:
        ctx.step(3)
        st = _by_id(ctx.download())
        ctx.step(1)
        d = ctx.download()
    ids = d["ids"].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(sc.n))
    nxt = _by_id(d)
    p = orc32.particles(st["pos"], st["vel"], ids=np.arange(sc.n))
    table = orc32.prepare(p, s_o)
    sorted_ids = p["acceleration"][:, 0].astype(np.int64)
    where = np.zeros(sc.n, dtype=np.int64)
    where[sorted_ids] = np.arange(sc.n)
    sample = np.random.default_rng(1).choice(sc.n, size=4000, replace=False)
    rho, cnt = orc32.density_subset(p, table, s_o, where[sample])
    orc32.free_table(table)
    rel = np.abs(nxt["rho"][sample].astype(np.float64) - rho) / rho
    assert rel.max() <= 1e-5, rel.max()
    assert cnt.mean() > 10


def test_fine_grid_scene_alone_equals_batched(sph):
    """Scenes on grids of reach 1, 2 and 3 in one batch: each is
    bit-identical to the same scene stepped alone on its own grid."""
    scenes = sph_scenes.batched_workload(nscenes=3, seed=5) + [sph_scenes.workload(0)]
    scales = [0.5, 1.0, 1.0 / 3.0, 0.5]
    descs = [sph.scene_desc(sc.n, sph.settings(*sc.settings_args()),
                            sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt, cell_scale=cs)
             for sc, cs in zip(scenes, scales)]
    with sph.Batch(descs) as b:
        for k, sc in enumerate(scenes):
            b.upload(k, sc.pos, sc.vel)
        b.step(40)
        mixed = [_bytes(b.download(k)) for k in range(len(scenes))]
    for k, (sc, cs) in enumerate(zip(scenes, scales)):
        with _ctx(sph, sc, cs) as ctx:
            ctx.step(40)
            assert _bytes(ctx.download()) == mixed[k], f"scene {k} ({sc.name}, scale {cs:.3f})"


def test_autotune_grid_restores_state_and_keeps_the_chosen_grid(sph):
    """The grid tuner tries block sizes x cell scales {1, 1/2}, restores the
    state, and the steps that follow are bit-identical to a fresh context
    (or, per scene, a batch-free context) on the chosen grid."""
    sc = sph_scenes.workload(0)
    with _ctx(sph, sc, 1.0) as ctx:
        t = ctx.autotune(3, grid=True)
        assert t["trials"] == 6 and t["cell_scale"] in (1.0, 0.5), t
        assert t["stencil_block"] in (64, 128, 256) and min(t["trial_ms"]) == t["ms_per_step"] > 0
        ctx.step(20)
        tuned = _bytes(ctx.download())
        ctx.upload(sc.pos, sc.vel)  # the chosen grid outlives an upload
        ctx.step(20)
        again = _bytes(ctx.download())
    assert again == tuned
    with _ctx(sph, sc, t["cell_scale"]) as ctx:
        ctx.step(20)
        assert _bytes(ctx.download()) == tuned

    scenes = sph_scenes.batched_workload(nscenes=3, seed=8)
    descs = [sph.scene_desc(s.n, sph.settings(*s.settings_args()),
                            sph.box(s.box_lo, s.box_hi, s.restitution, s.offset), s.dt) for s in scenes]
    with sph.Batch(descs) as b:
        for k, s in enumerate(scenes):
            b.upload(k, s.pos, s.vel)
        t = b.autotune(2, grid=True)
        assert t["trials"] == 6 and t["cell_scale"] in (1.0, 0.5), t
        b.step(10)
        mixed = [_bytes(b.download(k)) for k in range(len(scenes))]
    for k, s in enumerate(scenes):
        with _ctx(sph, s, t["cell_scale"]) as ctx:
            ctx.step(10)
            assert _bytes(ctx.download()) == mixed[k], f"scene {k}"


def test_auto_grid_follows_the_state_and_keeps_batch_identity(sph, orc32):
    """SPH_SCENE_AUTO_GRID: a jittered lattice (uniform at the scale of h)
    moves to cells of h / 2 after its first step while a clumped scene
    stays on cells of h (premises, from the sampled candidates per
    neighbour); over 70 steps (re-grid points at steps 0 and 64) each scene
    is bit-identical alone and in a batch beside scenes without the flag; a
    step from the lattice's state on its fine grid (step 40) matches the
    oracle fed the same state."""
    h = 0.03  # a 52^3 jittered lattice of spacing h / 2 (the dam break's start), above the size gate
    pos = sph_scenes.lattice((52, 52, 52), 0.5 * h, [-0.5, -0.9, -0.5], 0.01 * h, 3)
    lattice = sph_scenes.Scene("lattice-52", pos, np.zeros_like(pos), 1.6e-4, 1000.0,
                               1.0, 1.04, h, -9.8, 0.2, 0.001, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    assert lattice.n >= 1 << 17
    clumps = sph_scenes.blobs(nblobs=4, per_blob=2500)
    other = sph_scenes.batched_workload(nscenes=2, seed=9)
    scenes = [lattice, clumps] + other
    auto = [True, True, False, False]

    def desc(sc, a, keep_force=False):
        return sph.scene_desc(sc.n, sph.settings(*sc.settings_args()),
                              sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt,
                              auto_grid=a, keep_force=keep_force)

    with sph.Batch([desc(sc, a) for sc, a in zip(scenes, auto)]) as b:
        for k, sc in enumerate(scenes):
            b.upload(k, sc.pos, sc.vel)
        b.step(1)
        assert b.grid_info(0)[0] == 0.5, b.grid_info(0)
        assert b.grid_info(1)[0] == 1.0, b.grid_info(1)
        b.step(69)
        mixed = [_bytes(b.download(k)) for k in range(2)]
    for k in range(2):
        sc = scenes[k]
        with sph.Context(sc.n, sph.settings(*sc.settings_args()),
                         sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt, auto_grid=True) as ctx:
            ctx.upload(sc.pos, sc.vel)
            ctx.step(70)
            assert _bytes(ctx.download()) == mixed[k], f"scene {k} ({sc.name}) differs alone"
    # parity of the fine grid the lattice chose: one step from its state
    sc = lattice
    with sph.Context(sc.n, sph.settings(*sc.settings_args()),
                     sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt, auto_grid=True,
                     keep_force=True) as ctx:
        ctx.upload(sc.pos, sc.vel)
        ctx.step(40)
        assert ctx.grid_info()[0] == 0.5
        st = _by_id(ctx.download())
        ctx.step(1)
        p, _ = ctx.download_particles(transforms=False)
        d = ctx.download()
    nxt = _by_id(d)
    f = np.zeros((sc.n, 3), dtype=F32)
    f[d["ids"].astype(np.int64)] = p["force"][:, :3]
    s_o = orc32.settings(*sc.settings_args())
    box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    ref, _, rows, _, _ = pu.oracle_step(orc32, st["pos"], st["vel"], s_o, sc.dt, box_o)
    pu.assert_rho_p(nxt["rho"], nxt["pressure"], ref["density"], ref["pressure"], s_o, "auto grid")
    scale = pu.force_scale(st["pos"], st["vel"], ref["density"], ref["pressure"], rows, s_o)
    pu.assert_force(f, ref["force"][:, :3].astype(np.float64), scale, "auto grid")
