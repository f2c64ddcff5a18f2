"""CPU tests of libsph_hip.so: it loads, exports every symbol the header
declares, mirrors the reference structs and constants, and its per-particle
classification (the rule that lets the dense-grid kernels reproduce the
reference's 16-bit bucket multiplicities) holds against brute force.  No
device work is done here."""
import ctypes
import os
import re

import numpy as np
import pytest

import brute

F32 = np.float32
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "sph_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(sph):
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(sph.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_struct_sizes(sph):
    assert sph.load().sph_abi_version() == 1
    assert sph.PARTICLE_DTYPE.itemsize == 60
    assert ctypes.sizeof(sph.Settings) == 116
    assert ctypes.sizeof(sph.Box) == 32


@pytest.mark.parametrize("args", [(0.02, 1000.0, 1.0, 1.04, 0.15, -9.8, 0.2),
                                  (7.4074074e-4, 1000.0, 1.0, 1.04, 0.05, -9.8, 0.2),
                                  (1.6e-4, 1000.0, 1.0, 1.04, 0.03, -9.8, 0.2),
                                  (0.5, 800.0, 3.0, 0.5, 0.7, -1.0, 1.0)])
def test_settings_bitwise_equal_to_oracle(sph, orc32, args):
    a = sph.settings(*args)
    b = orc32.settings(*args)
    assert bytes(a) == bytes(b)


def test_box_presets(sph, orc32):
    a = sph.box_reference_cpu()
    b = orc32.box_reference_cpu()
    assert bytes(a) == bytes(b)
    g = sph.box_reference_gpu()
    assert list(g.lo) == [-10, 0, -10] and g.hi[0] == 10 and g.restitution == 1.0
    assert g.offset == F32(1e-5)


def test_hash_matches_oracle(sph, orc32):
    rng = np.random.default_rng(0)
    for c in rng.integers(-100000, 100000, size=(200, 3)):
        assert sph.get_hash(*map(int, c)) == orc32.get_hash(*map(int, c))


def test_classify_matches_reference_cell_and_bucket(sph, orc32):
    rng = np.random.default_rng(1)
    h = 0.15
    pos = rng.uniform(-3, 3, size=(500, 3)).astype(F32)
    for p in pos:
        h16, _, cell = sph.classify_host(p, h)
        assert cell == orc32.get_cell(tuple(map(float, p)), h)
        assert h16 == orc32.get_hash(*cell) & 0xFFFF


def test_slow_flag_covers_every_bucket_collision(sph):
    """buckets_collide (GF(2) rank test) agrees with a brute-force check of
    the 27 buckets; particles in colliding cells are always slow."""
    h = 1.0
    rng = np.random.default_rng(2)
    cells = np.concatenate([rng.integers(-300, 300, size=(4000, 3)),
                            np.array([[x, 0, 0] for x in range(-20, 20)] + [[0, y, 0] for y in range(-20, 20)] +
                                     [[0, 0, z] for z in range(-20, 20)] + [[5, 0, 7], [0, 0, 0]])])
    hits = 0
    for c in cells:
        # centre of the reference cell (cell 0 spans (-h, h))
        centre = [(ci + 0.5 * np.sign(ci)) * h if ci != 0 else 0.0 for ci in c]
        _, slow, cell = sph.classify_host(centre, h)
        assert cell == tuple(int(v) for v in c)
        collide = brute.buckets_collide_brute(c)
        hits += collide
        if collide:
            assert slow, c
        else:
            assert not slow, c  # the centre is far from boundaries
    assert hits > 50


def _eta(q):
    return 2.0 ** -21 + 2.0 ** -23 * (abs(q) + 1.0)


def test_near_boundary_particles_are_slow(sph):
    """Positions whose quotient x/h lies within eta(q) = 2^-21 + 2^-23 (|q|+1)
    of a nonzero integer are on the exact-multiplicity path; positions well
    outside that band (same row, same bucket structure) are not."""
    h = F32(0.05)
    y, z = F32(0.61) * h + F32(3.3), F32(5.43)
    for k in (1, 2, 7, -1, -3, 100, 1000):
        for frac in (0.0, 0.25, 0.9):
            for s in (-1, 1):
                q = F32(k + s * frac * _eta(k))
                x = F32(q * h)
                qx = float(x / h)
                if abs(qx - k) >= _eta(k):
                    continue  # rounding moved it out of the band
                _, slow, _ = sph.classify_host((x, y, z), float(h))
                assert slow, (k, frac, s)
        # far from the boundary: only a bucket collision could make it slow
        x = F32((k + 0.5) * h)
        _, slow, cell = sph.classify_host((x, y, z), float(h))
        assert slow == brute.buckets_collide_brute(cell), (k, cell)


def _fast_path_holds(sph, pos, h):
    """For every particle the classifier calls fast, every brute-force
    neighbour has multiplicity exactly 1."""
    rows = brute.neighbors_brute(pos, h)
    n_fast = 0
    for i, row in enumerate(rows):
        _, slow, _ = sph.classify_host(pos[i], h)
        if not slow:
            n_fast += 1
            assert len(row) == len(np.unique(row)), i
            # and no neighbour is missing: brute force counts every j with
            # fl(d2) < h2 whose bucket is among the 27 (m >= 1)
            d2 = brute.dist2_rows(pos, i)
            within = np.nonzero(d2 < F32(h) * F32(h))[0]
            within = within[within != i]
            assert np.array_equal(np.sort(within), row), i
    return n_fast


def test_fast_path_theorem_random_cloud(sph):
    h = 0.15
    rng = np.random.default_rng(3)
    pos = rng.uniform(-3 * h, 3 * h, size=(1500, 3)).astype(F32)
    assert _fast_path_holds(sph, pos, h) > 500


def test_fast_path_theorem_boundary_adversarial(sph):
    """Pairs straddling reference cell boundaries at distance just below h."""
    h = F32(0.05)
    rng = np.random.default_rng(4)
    pts = []
    for _ in range(300):
        k = int(rng.integers(-40, 40))
        b = F32(k) * h
        y, z = rng.uniform(-1, 1, size=2).astype(F32)
        dx = F32(rng.uniform(0.99, 1.0)) * h
        x0 = F32(b - F32(rng.uniform(0, 2e-6)) * rng.choice([1e-3, 1.0, 30.0]))
        pts.append((x0, y, z))
        pts.append((F32(x0 + dx), y, z))
    pos = np.array(pts, dtype=F32)
    _fast_path_holds(sph, pos, float(h))


@pytest.mark.parametrize("reach", [1, 2, 3])
def test_grid_cells_of_neighbours_are_adjacent(reach):
    """Completeness of the (2R + 1)^3 grid stencil: with cell edge
    H = h (1 + 2^-10) / R and double-precision grid coordinates, any pair
    with fl(d^2) < fl(h*h) lies at most R cells apart per axis, including
    pairs at distance h - ulp across a grid boundary."""
    h = F32(0.03)
    H = float(h) * (1.0 + 2.0 ** -10) / reach
    origin = np.array([-2.0 - H, -1.0 - H, -1.0 - H])
    inv = 1.0 / H
    dims = (140 * reach, 70 * reach, 70 * reach)
    rng = np.random.default_rng(5)
    pairs = []
    for _ in range(20000):
        g = rng.integers(1, 60)
        a = F32(origin[0] + g * H) - F32(rng.uniform(0, 1e-6))
        d = F32(rng.uniform(0.999, 1.0)) * h
        y, z = rng.uniform(-0.9, 0.9, size=2).astype(F32)
        pairs.append(((a, y, z), (F32(a + d), y, z)))
    p = np.array([q for pr in pairs for q in pr], dtype=F32)
    gc = brute.grid_coords(p, origin, inv, dims)
    h2 = h * h
    checked = 0
    for k in range(len(pairs)):
        i, j = 2 * k, 2 * k + 1
        dv = p[j] - p[i]
        d2 = (dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]
        if d2 < h2:
            checked += 1
            assert np.all(np.abs(gc[i] - gc[j]) <= reach)
    assert checked > 1000


@pytest.mark.parametrize("reach", [1, 2, 3])
def test_stencil_cut_keeps_every_neighbour(reach):
    """The per-lane stencil cut's margin: a lane at fraction f of its cell
    skips row (dy, dz) when gy^2 + gz^2 >= r2 and keeps the x-cells
    [floor(fx - w), floor(fx + w)] with w = sqrt(r2 - gy^2 - gz^2) +
    2^-11, where gap(d) = d - f (d > 0), -d - 1 + f (d < 0) or 0, less
    2^-12, r2 = (h / H)^2 (1 + 2^-10), all in float32 as the kernels
    evaluate them.  For 60,000 pairs with fl(d^2) < fl(h*h), the first
    placed just inside a random cell corner and the second just under h
    away in a random direction, the first's cell is never cut from the
    second's stencil."""
    h = F32(0.03)
    H = float(h) * (1.0 + 2.0 ** -10) / reach
    origin = np.array([-2.0 - H, -1.0 - H, -1.0 - H])
    inv = 1.0 / H
    dims = (140 * reach, 70 * reach, 70 * reach)
    rng = np.random.default_rng(17 + reach)
    m = 60000
    k = np.stack([rng.integers(reach + 2, 40 * reach, size=m) for _ in range(3)], axis=1)
    j = (origin + k * H + rng.uniform(0, 2e-7, size=(m, 3)) * rng.choice([0.0, 1.0], size=(m, 3))).astype(F32)
    u = rng.normal(size=(m, 3))
    u[rng.random((m, 3)) < 0.15] = 0.0
    u[~u.any(axis=1), 0] = 1.0
    u /= np.linalg.norm(u, axis=1)[:, None]
    i = (j.astype(np.float64) + float(h) * (1.0 - rng.uniform(0, 4e-7, size=(m, 1))) * u).astype(F32)
    dv = j - i
    d2 = (dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1]) + dv[:, 2] * dv[:, 2]
    nb = d2 < h * h
    assert nb.sum() > m // 4
    gi, gj = brute.grid_coords(i, origin, inv, dims), brute.grid_coords(j, origin, inv, dims)
    off = gj - gi
    assert np.all(np.abs(off[nb]) <= reach)
    f = np.clip((((i.astype(np.float64) - origin) * inv) - gi).astype(F32), F32(0), F32(1))
    slack = F32(2.0 ** -12)
    r2 = F32((float(h) * inv) ** 2 * (1.0 + 2.0 ** -10))

    def gap(d, fr):
        g = np.where(d > 0, d.astype(F32) - fr, np.where(d < 0, (-d - 1).astype(F32) + fr, F32(0)))
        return np.maximum(g - slack, F32(0)).astype(F32)

    ay, az = gap(off[:, 1], f[:, 1]), gap(off[:, 2], f[:, 2])
    rem = (r2 - (ay * ay + az * az)).astype(F32)
    kept_row = rem > 0
    w = (np.sqrt(np.maximum(rem, F32(0))) + F32(2) * slack).astype(F32)
    lo = np.maximum(np.floor(f[:, 0] - w).astype(np.int64), -reach)
    hi = np.minimum(np.floor(f[:, 0] + w).astype(np.int64), reach)
    kept = kept_row & (off[:, 0] >= lo) & (off[:, 0] <= hi)
    assert np.all(kept[nb]), f"{(~kept[nb]).sum()} neighbours cut"
    # the pairs sit at the margin: measuring the gap from the far side of
    # the lane's cell (d - 1 + f for d > 0) would cut neighbours
    bad = np.where(off[:, 1] > 0, np.maximum((off[:, 1] - 1).astype(F32) + f[:, 1] - slack, F32(0)), ay)
    assert ((r2 - (bad * bad + az * az)) <= 0)[nb].any() or (off[nb, 1] <= 0).all()
    # and the cut does cut: of the (2R + 1)^3 cells around lanes placed
    # uniformly in their cell, about a quarter (R = 1) to a third (R = 3)
    # are dropped
    fr = rng.uniform(0, 1, size=(4000, 3)).astype(F32)
    o = np.arange(-reach, reach + 1)
    dx, dy, dz = [a.ravel() for a in np.meshgrid(o, o, o, indexing="ij")]
    gx = gap(dx[None, :], fr[:, :1])
    g2 = gx * gx + gap(dy[None, :], fr[:, 1:2]) ** 2 + gap(dz[None, :], fr[:, 2:3]) ** 2
    assert (g2 < r2).mean() < [0.78, 0.69, 0.65][reach - 1]


def test_invalid_arguments_fail_without_touching_a_device(sph):
    lib = sph.load()
    s = sph.settings()
    assert lib.sph_hip_update_particles(None, None, 5, ctypes.byref(s), 0.003) == sph.SPH_E_INVALID
    assert b"NULL" in lib.sph_hip_last_error()
    assert lib.sph_hip_update_particles(None, None, 0, ctypes.byref(s), 0.003) == sph.SPH_OK
    bad = sph.settings(h=0.15)
    bad.h = 0.0
    p = sph.make_particles(np.zeros((3, 3), F32))
    assert lib.sph_hip_update_particles(p.ctypes.data_as(ctypes.c_void_p), None, 3, ctypes.byref(bad),
                                        0.003) == sph.SPH_E_INVALID
    d = sph.SceneDesc()
    d.n = 0
    h = ctypes.c_void_p()
    assert lib.sph_hip_create(ctypes.byref(d), ctypes.byref(h)) == sph.SPH_E_INVALID
    assert lib.sph_hip_step(None, 1) == sph.SPH_E_INVALID


@pytest.mark.parametrize("scale", [0.3, 0.25, -1.0, float("nan"), float("inf")])
def test_invalid_cell_scale_rejected(sph, scale):
    """Cell scales other than 0, >= 1 or [1/3, 1) are rejected before any
    device work."""
    sc = sph.settings(h=0.15)
    d = sph.scene_desc(5, sc, sph.box((-1, -1, -1), (1, 1, 1)), 0.003, cell_scale=scale)
    h = ctypes.c_void_p()
    assert sph.load().sph_hip_create(ctypes.byref(d), ctypes.byref(h)) == sph.SPH_E_INVALID
    assert b"cell_scale" in sph.load().sph_hip_last_error()


def test_server_structs_match_the_header(sph):
    # uint64 + 4 x uint32; 8 x uint64 + 7 x double + 2 x uint64
    assert ctypes.sizeof(sph.ServerConfig) == 24
    assert ctypes.sizeof(sph.ServerStats) == 136


def _glibc_rand(seed, count):
    """glibc rand() (random() of TYPE_3, the default state) restated from its
    published algorithm: r0 = seed; r_i = 16807 r_(i-1) mod (2^31 - 1) for
    i = 1..30; r_i = r_(i-31) for i = 31..33; r_i = r_(i-31) + r_(i-3)
    mod 2^32 after that; the k-th output is r_(k+344) >> 1."""
    r = [seed]
    for i in range(1, 31):
        r.append((16807 * r[i - 1]) % 2147483647)
    for i in range(31, 34):
        r.append(r[i - 31])
    for i in range(34, 344 + count):
        r.append((r[i - 31] + r[i - 3]) & 0xFFFFFFFF)
    return [v >> 1 for v in r[344:344 + count]]


def test_reference_initial_state_pinned_by_a_restated_glibc_rand():
    """sph_scenes.reference_init (SPHSystem::initParticles, src/SPHSystem.cpp:
    76-108, on libc's rand) equals the same state built on the restated
    glibc generator, whose first outputs for seed 1 are the well-known
    1804289383, 846930886, 1681692777."""
    import sph_scenes
    assert _glibc_rand(1, 3) == [1804289383, 846930886, 1681692777]
    w, h = 9, F32(0.15)
    seq = iter(_glibc_rand(1024, 3 * w ** 3))
    pos = np.zeros((w ** 3, 3), dtype=F32)
    sep = h + F32(0.01)
    for i in range(w):
        for j in range(w):
            for k in range(w):
                r = [(F32(next(seq)) / F32(2147483647) * F32(0.5) - F32(1)) * h / F32(10) for _ in range(3)]
                pos[i + (j + w * k) * w] = (F32(i) * sep + r[0] - F32(1.5), F32(j) * sep + r[1] + h + F32(0.1),
                                            F32(k) * sep + r[2] - F32(1.5))
    assert np.array_equal(pos, sph_scenes.reference_init(w, 0.15))
