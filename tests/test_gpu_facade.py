"""GPU tests of the C++ SPHSystem facade (the reference's class interface,
src/SPHSystem.h:24-62) through the headless step loop built on it: the
particles and render transforms equal the same scene stepped through the
resident C ABI, in the reference's order (sorted by 16-bit hash, ties by
id); reset() restores the initial state and closes the start gate; and a
facade step matches the oracle fed the state one step earlier."""
import os
import subprocess

import numpy as np
import pytest

import parity_util as pu
import sph_scenes

pytestmark = pytest.mark.gpu

HEADLESS = os.path.join(os.path.dirname(__file__), "..", "sph-fluid-simulator_amd", "bin", "sph_headless")


def _run(tmp_path, name, *args):
    out = tmp_path / name
    cmd = [HEADLESS, *[str(a) for a in args], "--dump", str(out)]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out


def _read_dump(path, dtype):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), np.uint64)[0])
        parts = np.frombuffer(f.read(60 * n), dtype=dtype)
        xf = np.frombuffer(f.read(64 * n), np.float32).reshape(n, 16)
        ids = np.frombuffer(f.read(4 * n), np.uint32)
    return parts, xf, ids


@pytest.mark.parametrize("lazy", [False, True])
def test_headless_facade_equals_resident_api(sph, tmp_path, lazy):
    width, steps = 12, 30
    out = _run(tmp_path, "dump.bin", "--width", width, "--steps", steps, *(["--lazy"] if lazy else []))
    parts, xf, ids = _read_dump(out, sph.PARTICLE_DTYPE)

    sc = sph_scenes.reference_app_scene(width)
    with sph.Context(sc.n, sph.settings(*sc.settings_args()), sph.box_reference_cpu(), 0.003,
                     keep_force=True) as ctx:
        ctx.upload(sc.pos)
        ctx.step(steps)
        p2, xf2 = ctx.download_particles()
        ids2 = ctx.download()["ids"]
    assert len(parts) == sc.n == width ** 3
    # the reference's order: by 16-bit bucket, ties by id
    order = np.lexsort((ids2, p2["hash"]))
    assert np.array_equal(ids, ids2[order])
    assert np.all(np.diff(parts["hash"].astype(np.int64)) >= 0)
    assert parts.tobytes() == p2[order].tobytes()  # the app's accelerations are zero, as the facade keeps them
    assert xf.tobytes() == xf2[order].tobytes()
    # the render transform is translate(position) * scale(h / 2)
    assert np.array_equal(xf[:, 12:15], parts["position"])
    assert np.all(xf[:, 0] == np.float32(0.15) / np.float32(2))


def test_facade_reset_restores_the_start(sph, tmp_path):
    """Seven updates, reset(), one update with the gate closed (must do
    nothing), then twenty updates: bit-identical to twenty updates of a
    fresh system (src/SPHSystem.cpp:136-143)."""
    a = _read_dump(_run(tmp_path, "a.bin", "--width", 10, "--steps", 20, "--reset-after", 7), sph.PARTICLE_DTYPE)
    b = _read_dump(_run(tmp_path, "b.bin", "--width", 10, "--steps", 20), sph.PARTICLE_DTYPE)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()
    a0 = _read_dump(_run(tmp_path, "a0.bin", "--width", 10, "--steps", 0, "--reset-after", 7), sph.PARTICLE_DTYPE)
    init = sph_scenes.reference_init(10, 0.15)
    assert np.array_equal(a0[0]["position"], init)  # reset() rebuilt the reference's initial state


def test_facade_step_matches_oracle(sph, orc32, tmp_path):
    """The reference application's scene (width 12) after 79 updates -- the
    bottom layers have reached the floor and interact: the 80th facade
    update against the oracle fed the same state: density and pressure
    within 1e-5, force within the scaled bar, position and velocity
    bit-exact when the oracle integrates with the facade's force and
    density, transforms translate(position) * scale(h / 2)."""
    width = 12
    s0 = _read_dump(_run(tmp_path, "s79.bin", "--width", width, "--steps", 79), sph.PARTICLE_DTYPE)
    s1 = _read_dump(_run(tmp_path, "s80.bin", "--width", width, "--steps", 80), sph.PARTICLE_DTYPE)
    n = width ** 3
    st = {k: np.zeros((n, 3), np.float32) for k in ("position", "velocity")}
    for k in st:
        st[k][s0[2].astype(np.int64)] = s0[0][k]
    nx = np.zeros(n, dtype=s1[0].dtype)
    nx[s1[2].astype(np.int64)] = s1[0]
    sc = sph_scenes.reference_app_scene(width)
    s_o = orc32.settings(*sc.settings_args())
    box_o = orc32.box_reference_cpu()
    ref, xf_ref, rows, _, _ = pu.oracle_step(orc32, st["position"], st["velocity"], s_o, 0.003, box_o)
    assert sum(len(r) for r in rows) > n, "premise: the particles interact"
    pu.assert_rho_p(nx["density"], nx["pressure"], ref["density"], ref["pressure"], s_o, "facade")
    scale = pu.force_scale(st["position"], st["velocity"], ref["density"], ref["pressure"], rows, s_o)
    pu.assert_force(nx["force"], ref["force"].astype(np.float64), scale, "facade")
    assert np.array_equal(nx["hash"], ref["hash"])
    q = orc32.particles(st["position"], st["velocity"])
    q["force"] = nx["force"]
    q["density"] = nx["density"]
    xf_q = orc32.integrate(q, s_o, 0.003, box_o)
    assert np.array_equal(q["position"], nx["position"])
    assert np.array_equal(q["velocity"], nx["velocity"])
    xf = np.zeros((n, 16), np.float32)
    xf[s1[2].astype(np.int64)] = s1[1]
    assert np.array_equal(xf, xf_q)
