"""GPU parity of the stateless drop-in (sph_hip_update_particles, the
replacement of updateParticles(..., onGPU), src/sph.h:19-22) and of the
neighbour multisets, against the oracle on identical inputs."""
import numpy as np
import pytest

import parity_util as pu
import sph_scenes

pytestmark = pytest.mark.gpu
F32 = np.float32


def _settings_pair(sph, orc32, scene):
    return sph.settings(*scene.settings_args()), orc32.settings(*scene.settings_args())


def _boxes(sph, orc32, scene):
    return (sph.box(scene.box_lo, scene.box_hi, scene.restitution, scene.offset),
            orc32.box(scene.box_lo, scene.box_hi, scene.restitution, scene.offset))


def _evolve_oracle(orc32, scene, steps):
    """A mid-simulation state: the oracle run from the scene's start."""
    s = orc32.settings(*scene.settings_args())
    box = orc32.box(scene.box_lo, scene.box_hi, scene.restitution, scene.offset)
    p = orc32.particles(scene.pos, scene.vel, ids=np.arange(scene.n))
    for _ in range(steps):
        orc32.update(p, s, scene.dt, box, transforms=False)
    out = np.zeros(scene.n, dtype=p.dtype)
    out[p["acceleration"][:, 0].astype(np.int64)] = p
    return out["position"].copy(), out["velocity"].copy()


def _axis_cloud(n, h, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-3 * h, 3 * h, size=(n, 3)).astype(F32)
    vel = rng.normal(scale=0.3, size=(n, 3)).astype(F32)
    return pos, vel


def _boundary_pairs(h, seed, npairs=400):
    """Pairs at distance just below h straddling reference cell boundaries:
    exercises the exact-multiplicity path and the grid-stencil margin."""
    rng = np.random.default_rng(seed)
    h = F32(h)
    pts = []
    for _ in range(npairs):
        k = int(rng.integers(-8, 8))
        b = F32(k) * h
        y, z = rng.uniform(-0.3, 0.3, size=2).astype(F32)
        x0 = F32(b - F32(rng.uniform(0, 2e-6)))
        pts.append((x0, y, z))
        pts.append((F32(x0 + F32(rng.uniform(0.995, 1.0)) * h), y, z))
    return np.array(pts, dtype=F32)


def check_dropin(sph, orc32, pos, vel, args, dt, box_lo, box_hi, e=0.5, off=1e-4, label="",
                 reference_box=False):
    s_g = sph.settings(*args)
    s_o = orc32.settings(*args)
    if reference_box:
        box_o = orc32.box_reference_cpu()
        box_g = None
    else:
        box_o = orc32.box(box_lo, box_hi, e, off)
        box_g = sph.box(box_lo, box_hi, e, off)
    n = len(pos)
    ref, xf_ref, rows_ref, _, _ = pu.oracle_step(orc32, pos, vel, s_o, dt, box_o)

    inp = sph.make_particles(pos, vel)
    inp["acceleration"][:, 0] = np.arange(n, dtype=F32)
    inp["acceleration"][:, 1] = F32(7.5)
    inp["acceleration"][:, 2] = -np.arange(n, dtype=F32)
    out, xf = sph.update_particles(inp, s_g, dt, box_g)

    # returned sorted by hash, ties in input order; acceleration carried
    assert np.all(np.diff(out["hash"].astype(np.int64)) >= 0), label
    ids = out["acceleration"][:, 0].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(n))
    for hv in np.unique(out["hash"]):
        assert np.all(np.diff(ids[out["hash"] == hv]) > 0)
    assert np.all(out["acceleration"][:, 1] == F32(7.5))
    assert np.array_equal(out["acceleration"][:, 2], -ids.astype(F32))
    g, _ = pu.by_id(out)
    xfg = np.zeros_like(xf)
    xfg[ids] = xf

    # hash of the incoming positions: bit-exact
    assert np.array_equal(g["hash"], ref["hash"]), label
    # density / pressure / force within the parity tolerances
    pu.assert_rho_p(g["density"], g["pressure"], ref["density"], ref["pressure"], s_o, label)
    scale = pu.force_scale(pos, vel, ref["density"], ref["pressure"], rows_ref, s_o)
    ratio = pu.assert_force(g["force"], ref["force"].astype(np.float64), scale, label)

    # integrate + walls + transform: bit-exact when fed the GPU's force/rho
    q = orc32.particles(pos, vel)
    q["force"] = g["force"]
    q["density"] = g["density"]
    xf_q = orc32.integrate(q, s_o, dt, box_o)
    assert np.array_equal(q["position"], g["position"]), label
    assert np.array_equal(q["velocity"], g["velocity"]), label
    assert np.array_equal(xf_q, xfg), label
    return ratio


def check_neighbors(sph, orc32, pos, args, label=""):
    s_g = sph.settings(*args)
    s_o = orc32.settings(*args)
    n = len(pos)
    p = orc32.particles(pos, ids=np.arange(n))
    table = orc32.prepare(p, s_o)
    off, idx = orc32.neighbors(p, table, s_o)
    orc32.free_table(table)
    ids = p["acceleration"][:, 0].astype(np.int64)
    want = [None] * n
    for si in range(n):
        want[ids[si]] = np.sort(ids[idx[off[si]:off[si + 1]]])
    goff, gidx = sph.neighbors(sph.make_particles(pos), s_g)
    total = 0
    multi = 0
    for i in range(n):
        row = gidx[goff[i]:goff[i + 1]].astype(np.int64)
        assert np.array_equal(row, want[i]), f"{label}: particle {i}"
        total += len(row)
        multi += int(len(row) != len(np.unique(row)))
    return total, multi


# ------------------------------------------------------------- neighbours

def test_neighbors_workload0(sph, orc32):
    sc = sph_scenes.workload(0)
    total, _ = check_neighbors(sph, orc32, sc.pos, sc.settings_args(), "workload0")
    assert total > 20 * sc.n


def test_neighbors_workload0_evolved(sph, orc32):
    sc = sph_scenes.workload(0)
    pos, _ = _evolve_oracle(orc32, sc, 60)
    check_neighbors(sph, orc32, pos, sc.settings_args(), "workload0@60")


def test_neighbors_axis_columns_double_count(sph, orc32):
    pos, _ = _axis_cloud(6000, 0.15, 21)
    _, multi = check_neighbors(sph, orc32, pos, (0.02, 1000.0, 1.0, 1.04, 0.15, -9.8, 0.2), "axis")
    assert multi > 0


def test_neighbors_boundary_pairs(sph, orc32):
    pos = _boundary_pairs(0.05, 22)
    total, _ = check_neighbors(sph, orc32, pos, (0.02, 1000.0, 1.0, 1.04, 0.05, -9.8, 0.2), "boundary")
    assert total > 0


def test_neighbors_workload1(sph, orc32):
    sc = sph_scenes.workload(1)
    check_neighbors(sph, orc32, sc.pos, sc.settings_args(), "workload1")


# --------------------------------------------------------------- drop-in

def test_dropin_workload0_start(sph, orc32):
    sc = sph_scenes.workload(0)
    check_dropin(sph, orc32, sc.pos, sc.vel, sc.settings_args(), sc.dt, sc.box_lo, sc.box_hi, label="w0@0")


def test_dropin_workload0_evolved(sph, orc32):
    sc = sph_scenes.workload(0)
    pos, vel = _evolve_oracle(orc32, sc, 80)
    check_dropin(sph, orc32, pos, vel, sc.settings_args(), sc.dt, sc.box_lo, sc.box_hi, label="w0@80")


def test_dropin_reference_app_default(sph, orc32):
    """SPHSystem(15, SPHSettings(0.02, 1000, 1, 1.04, 0.15, -9.8, 0.2)) in the
    reference CPU box, after the cube has fallen onto the floor."""
    sc = sph_scenes.reference_app_scene(15)
    pos, vel = _evolve_oracle(orc32, sc, 250)
    check_dropin(sph, orc32, pos, vel, sc.settings_args(), 0.003, None, None, label="app@250",
                 reference_box=True)


def test_dropin_axis_cloud(sph, orc32):
    pos, vel = _axis_cloud(5000, 0.15, 31)
    check_dropin(sph, orc32, pos, vel, (0.02, 1000.0, 1.0, 1.04, 0.15, -9.8, 0.2), 0.003,
                 (-1, -1, -1), (1, 1, 1), label="axis")


def test_dropin_boundary_pairs(sph, orc32):
    pos = _boundary_pairs(0.05, 32)
    vel = np.zeros_like(pos)
    check_dropin(sph, orc32, pos, vel, (7.4e-4, 1000.0, 1.0, 1.04, 0.05, -9.8, 0.2), 0.003,
                 (-2, -2, -2), (2, 2, 2), label="boundary")


def test_dropin_workload1(sph, orc32):
    sc = sph_scenes.workload(1)
    check_dropin(sph, orc32, sc.pos, sc.vel, sc.settings_args(), sc.dt, sc.box_lo, sc.box_hi, label="w1@0")


def test_dropin_edge_cases(sph, orc32):
    s = sph.settings()
    # empty input is a no-op
    out, xf = sph.update_particles(np.zeros(0, dtype=sph.PARTICLE_DTYPE), s, 0.003)
    assert len(out) == 0
    # one particle: free fall in the reference box
    check_dropin(sph, orc32, np.array([[0.1, 2.0, -0.3]], F32), np.zeros((1, 3), F32),
                 (0.02, 1000.0, 1.0, 1.04, 0.15, -9.8, 0.2), 0.003, None, None, reference_box=True,
                 label="single")
    # particles far apart outside the box (reflection) and below the floor
    pos = np.array([[-9.0, 0.5, 0.0], [9.5, 0.5, 3.0], [0.0, -0.4, 0.0], [3.0, 40.0, -8.5]], F32)
    vel = np.array([[1.0, 0, 0], [0, 0, 0], [0, -1.0, 0], [0, 0, 2.0]], F32)
    check_dropin(sph, orc32, pos, vel, (0.02, 1000.0, 1.0, 1.04, 0.15, -9.8, 0.2), 0.003, None, None,
                 reference_box=True, label="walls")


def test_dropin_coincident_particles_give_nan_like_reference(sph, orc32):
    # two particles at the same point: dist = 0, normalize(0) is NaN in the
    # reference, so its force is NaN; the GPU path must agree
    pos = np.array([[0.5, 1.0, 0.5], [0.5, 1.0, 0.5], [0.56, 1.0, 0.5]], F32)
    s_g = sph.settings()
    s_o = orc32.settings(0.02, 1000.0, 1.0, 1.04, 0.15, -9.8, 0.2)
    ref, _, _, _, _ = pu.oracle_step(orc32, pos, np.zeros_like(pos), s_o, 0.003, orc32.box_reference_cpu())
    inp = sph.make_particles(pos)
    inp["acceleration"][:, 0] = np.arange(3, dtype=F32)
    out, _ = sph.update_particles(inp, s_g, 0.003)
    g, _ = pu.by_id(out)
    assert np.array_equal(np.isnan(g["force"]), np.isnan(ref["force"]))
    assert np.isnan(ref["force"][0]).all()
    pu.assert_rho_p(g["density"], g["pressure"], ref["density"], ref["pressure"], s_o)


def test_dropin_repeatable(sph):
    sc = sph_scenes.workload(0)
    s = sph.settings(*sc.settings_args())
    b = sph.box(sc.box_lo, sc.box_hi)
    inp = sph.make_particles(sc.pos, sc.vel)
    a1, x1 = sph.update_particles(inp, s, sc.dt, b)
    a2, x2 = sph.update_particles(inp, s, sc.dt, b)
    assert a1.tobytes() == a2.tobytes() and x1.tobytes() == x2.tobytes()


def test_dropin_clumped_blobs(sph, orc32):
    """Cells of hundreds of particles: most waves take the wave-uniform
    candidate walk, some span several cells."""
    sc = sph_scenes.blobs()
    ratio = check_dropin(sph, orc32, sc.pos, sc.vel, sc.settings_args(), sc.dt, sc.box_lo, sc.box_hi,
                         sc.restitution, sc.offset, "blobs")
    assert ratio <= 1.0
