"""GPU tests of the resident API (state in HBM across steps): teacher-forced
parity with the oracle, determinism, graph replay, the renderer transform
buffer, the 1M-particle workload, and the 1000-step drift report."""
import json
import os

import numpy as np
import pytest

import parity_util as pu
import sph_scenes

pytestmark = pytest.mark.gpu
F32 = np.float32
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def make_ctx(sph, sc, keep_force=False):
    s = sph.settings(*sc.settings_args())
    b = sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    ctx = sph.Context(sc.n, s, b, sc.dt, keep_force=keep_force)
    ctx.upload(sc.pos, sc.vel)
    return ctx


def state_by_id(d):
    ids = d["ids"].astype(np.int64)
    out = {}
    for k in ("pos", "vel", "rho", "pressure"):
        a = np.zeros_like(d[k])
        a[ids] = d[k]
        out[k] = a
    return out


def test_resident_teacher_forced_steps(sph, orc32):
    """After k GPU steps, feed the GPU state to the oracle and compare the
    next GPU step: density/pressure within 1e-5, positions within the bound
    the force tolerance implies."""
    sc = sph_scenes.workload(0)
    s_o = orc32.settings(*sc.settings_args())
    box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    with make_ctx(sph, sc) as ctx:
        done = 0
        for k in (0, 20, 60):
            ctx.step(k - done)
            st = state_by_id(ctx.download())
            ctx.step(1)
            done = k + 1
            nxt = state_by_id(ctx.download())
            ref, _, rows, _, _ = pu.oracle_step(orc32, st["pos"], st["vel"], s_o, sc.dt, box_o)
            pu.assert_rho_p(nxt["rho"], nxt["pressure"], ref["density"], ref["pressure"], s_o, f"step {k}")
            scale = pu.force_scale(st["pos"], st["vel"], ref["density"], ref["pressure"], rows, s_o)
            tol_v = pu.RTOL * scale / ref["density"] * sc.dt + 4 * np.spacing(np.abs(ref["velocity"]).max(1))
            dv = np.linalg.norm(nxt["vel"].astype(np.float64) - ref["velocity"], axis=1)
            assert np.all(dv <= tol_v), f"step {k}: velocity"
            dx = np.linalg.norm(nxt["pos"].astype(np.float64) - ref["position"], axis=1)
            assert np.all(dx <= tol_v * sc.dt + 4 * np.spacing(np.abs(ref["position"]).max(1))), f"step {k}: pos"


def test_resident_deterministic_and_graph_equivalent(sph):
    sc = sph_scenes.workload(0)
    outs = []
    for graph in (True, True, False):
        with make_ctx(sph, sc) as ctx:
            ctx.set_graph(graph)
            ctx.step(150)
            d = ctx.download()
            outs.append(np.concatenate([d["pos"].ravel(), d["vel"].ravel(), d["rho"], d["pressure"]]).tobytes()
                        + d["ids"].tobytes())
    assert outs[0] == outs[1], "two runs differ"
    assert outs[0] == outs[2], "graph replay differs from direct launches"


def test_resident_transforms_and_particles(sph, orc32):
    sc = sph_scenes.workload(0)
    with make_ctx(sph, sc, keep_force=True) as ctx:
        ctx.step(5)
        p, xf = ctx.download_particles()
        d = ctx.download()
    assert np.array_equal(p["position"], d["pos"])
    assert np.array_equal(p["density"], d["rho"])
    s_o = orc32.settings(*sc.settings_args())
    # transform of each position: glm::translate(x) * scale(h / 2)
    q = orc32.particles(p["position"])
    q["density"] = 1.0
    box = orc32.box((-1e30,) * 3, (1e30,) * 3)
    xf_ref = orc32.integrate(q, s_o, 0.0, box)
    assert np.array_equal(xf, xf_ref)
    # hash field: bucket of the positions at the start of the last step
    assert np.all(p["hash"] < 65536) and np.any(p["force"] != 0)


def test_resident_workload1_steps(sph, orc32):
    sc = sph_scenes.workload(1)
    s_o = orc32.settings(*sc.settings_args())
    box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    with make_ctx(sph, sc) as ctx:
        ctx.step(10)
        st = state_by_id(ctx.download())
        ctx.step(1)
        nxt = state_by_id(ctx.download())
    ref, _, _, _, _ = pu.oracle_step(orc32, st["pos"], st["vel"], s_o, sc.dt, box_o)
    pu.assert_rho_p(nxt["rho"], nxt["pressure"], ref["density"], ref["pressure"], s_o, "w1")


def test_workload2_full_size_sampled(sph, orc32):
    """1,000,000 particles: one GPU step; densities of 4000 sampled particles
    against the oracle; ids stay a permutation; everything inside the box."""
    sc = sph_scenes.workload(2)
    s_o = orc32.settings(*sc.settings_args())
    with make_ctx(sph, sc) as ctx:
        ctx.step(3)
        st = state_by_id(ctx.download())
        ctx.step(1)
        d = ctx.download()
    ids = d["ids"].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(sc.n))
    nxt = state_by_id(d)
    assert np.isfinite(nxt["pos"]).all()
    lo = np.array(sc.box_lo) + sc.h
    hi = np.array(sc.box_hi) - sc.h
    assert np.all(nxt["pos"] >= lo - 0.01) and np.all(nxt["pos"] <= hi + 0.01)
    p = orc32.particles(st["pos"], st["vel"], ids=np.arange(sc.n))
    table = orc32.prepare(p, s_o)
    sorted_ids = p["acceleration"][:, 0].astype(np.int64)
    where = np.zeros(sc.n, dtype=np.int64)
    where[sorted_ids] = np.arange(sc.n)
    sample = np.random.default_rng(0).choice(sc.n, size=4000, replace=False)
    rho, cnt = orc32.density_subset(p, table, s_o, where[sample])
    orc32.free_table(table)
    rel = np.abs(nxt["rho"][sample].astype(np.float64) - rho) / rho
    assert rel.max() <= 1e-5, rel.max()
    assert cnt.mean() > 10


def test_drift_report_workload0(sph, orc32, orc64):
    """1000 free-running steps of workload 0: GPU FP32 and oracle FP32 each
    against the oracle in FP64 (same method).  The system is chaotic, so the
    contract is that the GPU drifts no faster than the reference's own FP32
    arithmetic.  The report is printed, and written to
    $SPH_REPORT_DIR/drift_workload0.json when that directory is set."""
    sc = sph_scenes.workload(0)
    checkpoints = [1, 10, 100, 1000]
    s32 = orc32.settings(*sc.settings_args())
    s64 = orc64.settings(*sc.settings_args())
    b32 = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    b64 = orc64.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    p32 = orc32.particles(sc.pos, sc.vel, ids=np.arange(sc.n))
    p64 = orc64.particles(sc.pos.astype(np.float64), sc.vel.astype(np.float64), ids=np.arange(sc.n))
    report = {"workload": sc.name, "steps": {}}
    done = 0
    with make_ctx(sph, sc) as ctx:
        for c in checkpoints:
            for _ in range(c - done):
                orc32.update(p32, s32, sc.dt, b32, transforms=False)
                orc64.update(p64, s64, sc.dt, b64, transforms=False)
            ctx.step(c - done)
            done = c
            g = state_by_id(ctx.download())["pos"].astype(np.float64)
            o32 = np.zeros((sc.n, 3))
            o32[p32["acceleration"][:, 0].astype(np.int64)] = p32["position"]
            o64 = np.zeros((sc.n, 3))
            o64[p64["acceleration"][:, 0].astype(np.int64)] = p64["position"]
            dg = np.linalg.norm(g - o64, axis=1)
            dc = np.linalg.norm(o32 - o64, axis=1)
            report["steps"][c] = {
                "gpu_fp32_vs_fp64": {"max": float(dg.max()), "rms": float(np.sqrt((dg ** 2).mean())),
                                     "p99": float(np.percentile(dg, 99))},
                "cpu_fp32_vs_fp64": {"max": float(dc.max()), "rms": float(np.sqrt((dc ** 2).mean())),
                                     "p99": float(np.percentile(dc, 99))},
            }
    out_dir = os.environ.get("SPH_REPORT_DIR", "")
    if out_dir and os.path.isdir(out_dir):
        with open(os.path.join(out_dir, "drift_workload0.json"), "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps(report))
    for c in (1, 10, 100):
        r = report["steps"][c]
        assert r["gpu_fp32_vs_fp64"]["rms"] <= 10 * r["cpu_fp32_vs_fp64"]["rms"] + 1e-6, (c, r)


def _blob_states(sph, sc, lane_walk, chunks=4, every=10):
    with sph.Context(sc.n, sph.settings(*sc.settings_args()),
                     sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt,
                     lane_walk=lane_walk) as ctx:
        ctx.upload(sc.pos, sc.vel)
        out = []
        for _ in range(chunks):
            ctx.step(every)
            out.append(ctx.download())
    return out


def test_union_walk_matches_lane_walk_unsplit(sph):
    """Where no wave task is split (every walk below the part size), the
    wave-uniform candidate walk and the per-lane walk give bit-identical
    states step after step (premise: union walks occur along the run and no
    task is split)."""
    sc = sph_scenes.workload(0)
    with make_ctx(sph, sc) as ctx:
        unions = 0
        for _ in range(3):
            ctx.step(15)
            dump = ctx.walk_dump()
            unions += int(dump["union"].sum())
            assert not dump["split"].any(), "premise: no task is split"
    assert unions > 0, "premise: some waves take the union walk"
    u = _blob_states(sph, sc, False, chunks=3, every=15)
    l = _blob_states(sph, sc, True, chunks=3, every=15)
    for k, (x, y) in enumerate(zip(u, l)):
        for key in ("pos", "vel", "rho", "pressure", "ids"):
            assert x[key].tobytes() == y[key].tobytes(), f"{key} differs after {15 * (k + 1)} steps"


def test_split_walks_deterministic_and_agree(sph, orc32):
    """Clumped scene with split wave tasks: repeated runs are bit-identical;
    the two walks agree within rounding; both match the oracle's densities."""
    sc = sph_scenes.blobs()
    a = _blob_states(sph, sc, False, chunks=2, every=5)
    b = _blob_states(sph, sc, False, chunks=2, every=5)
    for x, y in zip(a, b):
        for key in ("pos", "vel", "rho", "pressure", "ids"):
            assert x[key].tobytes() == y[key].tobytes()
    s_g = sph.settings(*sc.settings_args())
    res = []
    for lane_walk in (False, True):
        with sph.Context(sc.n, s_g, sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt,
                         lane_walk=lane_walk) as ctx:
            ctx.upload(sc.pos, sc.vel)
            ctx.step(1)
            d = ctx.download()
            st = {}
            for key in ("pos", "vel", "rho", "pressure"):
                st[key] = np.zeros_like(d[key])
                st[key][d["ids"].astype(np.int64)] = d[key]
            res.append(st)
    u, l = res
    s_o = orc32.settings(*sc.settings_args())
    ref, _, _, _, _ = pu.oracle_step(orc32, sc.pos, sc.vel, s_o, sc.dt,
                                     orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset))
    pu.assert_rho_p(u["rho"], u["pressure"], ref["density"], ref["pressure"], s_o, "union walk")
    pu.assert_rho_p(l["rho"], l["pressure"], ref["density"], ref["pressure"], s_o, "lane walk")
    dv_ref = np.abs(l["vel"] - sc.vel).max(axis=1)
    dv = np.abs(u["vel"] - l["vel"]).max(axis=1)
    assert np.all(dv <= 1e-3 * dv_ref + 1e-5), float(np.max(dv - 1e-3 * dv_ref))


def test_slow_lane_clump_parity_and_walk_identity(sph, orc32):
    """A clump with slow lanes (colliding buckets: m_ij = 2): waves holding
    a slow lane take the per-lane walk, which applies the exact
    multiplicities (premise: every wave of this clump holds one, so all
    walk per lane; their long walks are split into parts).  One step
    matches the oracle (density, pressure, force), and the run stays
    bit-identical to one with the per-lane walk forced."""
    import brute
    sc = pu.collision_clump()
    cells = brute.ref_cells(sc.pos, sc.h)
    slow = np.array([brute.buckets_collide_brute(c) for c in cells])
    assert slow.sum() >= 64, "premise: at least a wave's worth of slow particles"
    rows = brute.neighbors_brute(sc.pos, sc.h)
    assert any(len(r) != len(np.unique(r)) for r in rows), "premise: some m_ij = 2"

    s_o = orc32.settings(*sc.settings_args())
    box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    ref, _, orows, _, _ = pu.oracle_step(orc32, sc.pos, sc.vel, s_o, sc.dt, box_o)
    with make_ctx(sph, sc, keep_force=True) as ctx:
        dump = ctx.walk_dump()
        assert not dump["union"].any(), "premise: per-lane walks"
        ctx.step(1)
        p, _ = ctx.download_particles(transforms=False)
        d = ctx.download()
    ids = d["ids"].astype(np.int64)
    nxt = state_by_id(d)
    f = np.zeros((sc.n, 3), dtype=F32)
    f[ids] = p["force"][:, :3]
    pu.assert_rho_p(nxt["rho"], nxt["pressure"], ref["density"], ref["pressure"], s_o, "collision clump")
    scale = pu.force_scale(sc.pos, sc.vel, ref["density"], ref["pressure"], orows, s_o)
    pu.assert_force(f, ref["force"][:, :3].astype(np.float64), scale, "collision clump")

    u = _blob_states(sph, sc, False, chunks=3, every=2)
    l = _blob_states(sph, sc, True, chunks=3, every=2)
    for k, (x, y) in enumerate(zip(u, l)):
        for key in ("pos", "vel", "rho", "pressure", "ids"):
            assert x[key].tobytes() == y[key].tobytes(), f"{key} differs after {2 * (k + 1)} steps"
