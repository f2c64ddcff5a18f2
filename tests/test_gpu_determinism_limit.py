"""GPU tests of the determinism limit: the in-cell sort orders a cell's
particles by id only up to 16,384 particles; a larger cell keeps the
arrival order of the atomic scatter, and every such cell is counted
(sph_hip_unordered_cells).  A clump inside one grid cell (the box corner,
side 0.2 h, so it lies in the first cell whatever the cell size >= h/2)
holds exactly n particles in that cell."""
import numpy as np
import pytest

import sph_scenes

pytestmark = pytest.mark.gpu
F32 = np.float32
LIMIT = 16384


def _corner_clump(n, seed=5, h=0.05):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.1 * h, 0.3 * h, size=(n, 3)).astype(F32)
    vel = np.zeros((n, 3), F32)
    return sph_scenes.Scene(f"corner-clump-{n}", pos, vel, 1e-6, 1000.0, 1.0, 1.04, h, -9.8, 0.2, 1e-7,
                            (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), steps=2)


def _run(sph, sc, steps, graph=True):
    with sph.Context(sc.n, sph.settings(*sc.settings_args()),
                     sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt) as ctx:
        ctx.set_graph(graph)
        ctx.upload(sc.pos, sc.vel)
        ctx.step(steps)
        d = ctx.download()
        return d, ctx.unordered_cells()


def test_cell_at_the_limit_is_sorted_and_reproducible(sph):
    sc = _corner_clump(LIMIT)
    a, ua = _run(sph, sc, 2)
    b, ub = _run(sph, sc, 2, graph=False)
    assert ua == 0 and ub == 0
    for k in ("pos", "vel", "rho", "pressure", "ids"):
        assert a[k].tobytes() == b[k].tobytes(), k


def test_cell_above_the_limit_is_counted_and_within_parity(sph, orc32):
    """One cell of 16,385 particles: counted once per step; the densities
    of the first step still agree with the oracle within 1e-5 (only the
    summation order is open)."""
    sc = _corner_clump(LIMIT + 1)
    d, u = _run(sph, sc, 1)
    assert u == 1
    _, u2 = _run(sph, sc, 2)
    assert u2 == 2
    s_o = orc32.settings(*sc.settings_args())
    p = orc32.particles(sc.pos, sc.vel, ids=np.arange(sc.n))
    table = orc32.prepare(p, s_o)
    try:
        orc32.density(p, table, s_o)
    finally:
        orc32.free_table(table)
    rho_ref = np.zeros(sc.n)
    rho_ref[p["acceleration"][:, 0].astype(np.int64)] = p["density"]
    rho = np.zeros(sc.n)
    rho[d["ids"].astype(np.int64)] = d["rho"]
    assert np.all(np.abs(rho - rho_ref) <= 1e-5 * rho_ref)


def test_batch_counts_its_scenes(sph):
    small, big = _corner_clump(1000), _corner_clump(LIMIT + 100, seed=6)
    descs = [sph.scene_desc(sc.n, sph.settings(*sc.settings_args()),
                            sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt) for sc in (small, big)]
    with sph.Batch(descs) as b:
        for k, sc in enumerate((small, big)):
            b.upload(k, sc.pos, sc.vel)
        b.step(3)
        assert b.unordered_cells() == 3
