"""GPU parity of the walks a production step takes -- the part plan, the
wave-uniform union walk, split parts, the per-lane walk and the neighbour
lists the force pass reads -- against the oracle.  sph_hip_walk_dump runs
the step's own passes on a state and returns, per particle, the neighbour
count with multiplicity (the terms src/sph.cpp:40-66 sums), which walk its
wave task took and its neighbour list; the forces come from the step
itself (SPH_SCENE_KEEP_FORCE).  Counts and lists are bit-exact; forces are
held to the scaled parity bar."""
import numpy as np
import pytest

import parity_util as pu
import sph_scenes

pytestmark = pytest.mark.gpu
F32 = np.float32


def _ctx(sph, sc, pos, vel, keep_force=False):
    ctx = sph.Context(sc.n, sph.settings(*sc.settings_args()),
                      sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset), sc.dt, keep_force=keep_force)
    ctx.upload(pos, vel)
    return ctx


def _by_id(d):
    ids = d["ids"].astype(np.int64)
    out = {}
    for k in ("pos", "vel", "rho", "pressure"):
        a = np.zeros_like(d[k])
        a[ids] = d[k]
        out[k] = a
    return out


class OracleState:
    """The oracle's view of one state: sorted particles, table, and (on
    demand) the densities and pressures of every particle."""

    def __init__(self, orc, sc, pos, vel):
        self.orc = orc
        self.s = orc.settings(*sc.settings_args())
        self.p = orc.particles(pos, vel, ids=np.arange(len(pos)))
        self.table = orc.prepare(self.p, self.s)
        self.sorted_ids = self.p["acceleration"][:, 0].astype(np.int64)
        self.where = np.zeros(len(pos), dtype=np.int64)
        self.where[self.sorted_ids] = np.arange(len(pos))
        self.have_density = False

    def rows(self, ids):
        """Neighbour rows (particle ids, multiplicity-repeated) of ids."""
        off, idx = self.orc.neighbors_subset(self.p, self.table, self.s, self.where[ids])
        return [np.sort(self.sorted_ids[idx[off[a]:off[a + 1]]]) for a in range(len(ids))]

    def forces(self, ids):
        if not self.have_density:
            self.orc.density(self.p, self.table, self.s)
            self.have_density = True
        return self.orc.forces_subset(self.p, self.table, self.s, self.where[ids]).astype(np.float64)

    def rho_p_by_id(self):
        rho = np.zeros(len(self.p))
        prs = np.zeros(len(self.p))
        rho[self.sorted_ids] = self.p["density"]
        prs[self.sorted_ids] = self.p["pressure"]
        return rho, prs

    def close(self):
        self.orc.free_table(self.table)


def _check_dump(dump, rows, ids, label):
    cnt = np.array([len(r) for r in rows])
    bad = np.nonzero(dump["count"][ids] != cnt)[0]
    assert len(bad) == 0, f"{label}: {len(bad)} neighbour counts differ, first id {ids[bad[0]]}: " \
                          f"{dump['count'][ids[bad[0]]]} vs {cnt[bad[0]]}"
    nlisted = 0
    for a, i in enumerate(ids):
        if not dump["listed"][i]:
            assert dump["list_count"][i] == 0xFFFFFFFF
            continue
        uniq = np.unique(rows[a])
        k = int(dump["list_count"][i])
        assert k == len(uniq), f"{label}: id {i} lists {k} of {len(uniq)} neighbours"
        assert np.array_equal(np.sort(dump["list"][i, :k].astype(np.int64)), uniq), f"{label}: list of id {i}"
        nlisted += 1
    return nlisted


def _check_forces(sph, orc32, sc, pos, vel, ids, ost, label):
    """One step from (pos, vel) with forces kept; the forces of ids against
    the oracle's, within the scaled bar."""
    with _ctx(sph, sc, pos, vel, keep_force=True) as ctx:
        ctx.step(1)
        p, _ = ctx.download_particles(transforms=False)
        d = ctx.download()
    f = np.zeros((sc.n, 3), dtype=F32)
    f[d["ids"].astype(np.int64)] = p["force"][:, :3]
    f_ref = ost.forces(ids)
    rho, prs = ost.rho_p_by_id()
    rows = [np.zeros(0, np.int64)] * sc.n
    for a, r in zip(ids, ost.rows(ids)):
        rows[a] = r
    scale = pu.force_scale(pos, vel, rho, prs, rows, ost.s)[ids]
    return pu.assert_force(f[ids], f_ref, scale, label)


def test_walk_dump_matches_oracle_workload0(sph, orc32):
    """8,000-particle cube at its start and after 60 steps (oracle-evolved):
    every particle's neighbour count (with multiplicity) and neighbour list
    equal the oracle's rows; the lattice start is all per-lane walks with
    lists, so the force pass reads lists there; the forces match."""
    sc = sph_scenes.workload(0)
    s_o = orc32.settings(*sc.settings_args())
    box_o = orc32.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset)
    p = orc32.particles(sc.pos, sc.vel, ids=np.arange(sc.n))
    states = [(sc.pos.copy(), sc.vel.copy())]
    for _ in range(60):
        orc32.update(p, s_o, sc.dt, box_o, transforms=False)
    back = np.zeros(sc.n, dtype=p.dtype)
    back[p["acceleration"][:, 0].astype(np.int64)] = p
    states.append((back["position"].copy(), back["velocity"].copy()))
    ids = np.arange(sc.n)
    for k, (pos, vel) in enumerate(states):
        label = f"workload0 state {k}"
        with _ctx(sph, sc, pos, vel) as ctx:
            dump = ctx.walk_dump(lists=True)
        ost = OracleState(orc32, sc, pos, vel)
        try:
            nlisted = _check_dump(dump, ost.rows(ids), ids, label)
            if k == 0:
                assert nlisted == sc.n, f"{label}: premise, every lattice particle listed"
            _check_forces(sph, orc32, sc, pos, vel, ids, ost, label)
        finally:
            ost.close()


def test_walk_dump_matches_oracle_on_clumps(sph, orc32):
    """Gaussian clumps beside a sparse corner: union walks, split tasks and
    listed per-lane walks all occur (premise), and every particle's count
    and list equal the oracle's; the forces of the step match."""
    sc = sph_scenes.blobs()
    corner = np.random.default_rng(41).uniform(0.65, 0.95, size=(2000, 3)).astype(F32)
    sc.pos = np.concatenate([sc.pos, corner])
    sc.vel = np.concatenate([sc.vel, np.zeros_like(corner)])
    with _ctx(sph, sc, sc.pos, sc.vel) as ctx:
        dump = ctx.walk_dump(lists=True)
    assert dump["union"].any() and dump["split"].any() and dump["listed"].any(), "premise: all walks occur"
    ids = np.arange(sc.n)
    ost = OracleState(orc32, sc, sc.pos, sc.vel)
    try:
        _check_dump(dump, ost.rows(ids), ids, "blobs")
        _check_forces(sph, orc32, sc, sc.pos, sc.vel, ids, ost, "blobs")
    finally:
        ost.close()


def _stratified(rng, dump, n_each):
    """Up to n_each particle ids from each walk class present."""
    classes = [dump["union"] & ~dump["split"], dump["split"], dump["listed"],
               ~dump["union"] & ~dump["split"] & ~dump["listed"]]
    out = []
    for c in classes:
        idx = np.nonzero(c)[0]
        if len(idx):
            out.append(rng.choice(idx, size=min(n_each, len(idx)), replace=False))
    return np.unique(np.concatenate(out))


def test_walks_and_forces_at_1m_sampled(sph, orc32):
    """1,000,000-particle dam break at its start and after 600 GPU steps
    (clumped: union walks and split tasks occur, premise): the neighbour
    counts of >= 4,000 particles sampled across the walk classes are
    bit-exact, their lists equal the oracle's sets, and their forces of the
    next step are within the scaled bar."""
    sc = sph_scenes.workload(2)
    rng = np.random.default_rng(12)
    with _ctx(sph, sc, sc.pos, sc.vel) as ctx:
        d0 = ctx.download()
        ctx.step(600)
        d1 = ctx.download()
    for k, d in enumerate((d0, d1)):
        st = _by_id(d)
        label = f"1M state {k}"
        print(f"{label}: dump", flush=True)
        with _ctx(sph, sc, st["pos"], st["vel"]) as ctx:
            dump = ctx.walk_dump(lists=True)
        if k == 1:
            assert dump["union"].sum() > 0 and dump["split"].sum() > 0, "premise: clumped walks occur"
        # every walk class present, plus a uniform sample
        ids = np.unique(np.concatenate([_stratified(rng, dump, 1000), rng.choice(sc.n, size=4000, replace=False)]))
        print(f"{label}: oracle on {len(ids)} sampled particles", flush=True)
        ost = OracleState(orc32, sc, st["pos"], st["vel"])
        try:
            _check_dump(dump, ost.rows(ids), ids, label)
            _check_forces(sph, orc32, sc, st["pos"], st["vel"], ids, ost, label)
        finally:
            ost.close()


def _corner_pairs(h, scale, npairs, seed):
    """Pairs whose first particle sits just inside the low corner of a grid
    cell and whose second lies just under h from it in a random direction
    of the negative octant (some directions in a coordinate plane or along
    an axis), inside the box [-0.3, 0.3]^3.  The first particle's cell is
    then the farthest cell of the second's stencil that can hold a
    neighbour, the case the per-lane stencil cut decides at its margin.
    Returns positions and the grid (origin, edge, cells per axis) a scene
    of them gets."""
    H = h * (1.0 + 2.0 ** -10) * scale
    o = -0.3 - H
    dims = int(np.floor(0.6 / H)) + 3
    R = int(np.ceil(1.0 / scale - 1e-9))
    rng = np.random.default_rng(seed)
    pts = []
    for _ in range(npairs):
        k = rng.integers(R + 3, dims - R - 3, size=3)
        j = (o + k * H + 1e-6).astype(F32)
        u = rng.uniform(0.05, 1.0, size=3)
        u[rng.random(3) < 0.2] = 0.0
        if not u.any():
            u[0] = 1.0
        u /= np.linalg.norm(u)
        i = (j.astype(np.float64) - h * (1.0 - rng.uniform(0.0, 3e-7)) * u).astype(F32)
        pts += [j, i]
    return np.array(pts, dtype=F32), (o, H, dims)


@pytest.mark.parametrize("scale", [1.0, 0.5, 1.0 / 3.0])
@pytest.mark.parametrize("lane_walk", [True, False])
def test_stencil_cut_keeps_pairs_at_the_corner_margin(sph, orc32, scale, lane_walk):
    """The per-lane walk that records neighbour lists visits only the cells
    of its stencil within h of the lane (rows beyond h skipped, x-runs cut
    to the chord): pairs at distance just under h reaching exactly into a
    cell's corner are all found -- counts bit-exact and lists equal to the
    oracle's -- on grids of reach 1, 2 and 3, with every wave on the
    per-lane walk or with the plan's choice."""
    h = 0.05
    pos, (o, H, dims) = _corner_pairs(h, scale, 1500, seed=int(77 + 10 * scale))
    n = len(pos)
    assert dims ** 3 <= max(64 * n, 1 << 16), "premise: the grid is not coarsened"
    assert pos.min() > -0.3 and pos.max() < 0.3
    d = pos[0::2] - pos[1::2]
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    assert (d2 < F32(h) * F32(h)).sum() > 200, "premise: most pairs are neighbours"
    sc = sph_scenes.Scene("corner-pairs", pos, np.zeros_like(pos), 0.02 * (h / 0.15) ** 3, 1000.0, 1.0, 1.04,
                          h, -9.8, 0.2, 1e-4, (-0.3, -0.3, -0.3), (0.3, 0.3, 0.3), steps=1)
    with sph.Context(n, sph.settings(*sc.settings_args()), sph.box(sc.box_lo, sc.box_hi, sc.restitution, sc.offset),
                     sc.dt, cell_scale=scale, lane_walk=lane_walk) as ctx:
        ctx.upload(sc.pos, sc.vel)
        dump = ctx.walk_dump(lists=True)
    if lane_walk:
        assert not dump["union"].any()
    ids = np.arange(n)
    ost = OracleState(orc32, sc, sc.pos, sc.vel)
    try:
        _check_dump(dump, ost.rows(ids), ids, f"corner pairs, scale {scale:.3f}")
    finally:
        ost.close()
