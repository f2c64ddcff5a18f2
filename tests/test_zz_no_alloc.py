"""GPU test of sph_hip_update_particles_noalloc, the drop-in step whose
device buffers all come from one fixed array in device memory: called
through the C ABI twice in a row (the second call smaller, so it finds the
first call's contents in the array), its particles and transforms must be
bit-identical to sph_hip_update_particles_box on the same inputs.  Sizes
beyond SPH_NOALLOC_MAX_PARTICLES are rejected before any device work."""
import numpy as np
import pytest

import sph_scenes

pytestmark = pytest.mark.gpu


def _run_both(sph, scene):
    s = sph.settings(*scene.settings_args())
    box = sph.box(scene.box_lo, scene.box_hi, scene.restitution, scene.offset)
    p = sph.make_particles(scene.pos, scene.vel)
    p["acceleration"][:, 0] = np.arange(scene.n, dtype=np.float32)
    ref, xf_ref = sph.update_particles(p, s, scene.dt, box)
    out, xf = sph.update_particles_noalloc(p, s, scene.dt, box)
    return ref, xf_ref, out, xf


def test_noalloc_matches_allocating_dropin_on_two_calls(sph):
    # first call: 20,000 particles in dense clumps (split wave tasks, big
    # cells); second call: a smaller 6,000-particle cloud straddling the
    # axes (slow lanes with m_ij = 2) that lands on the first call's
    # leftovers in the array
    big = sph_scenes.blobs(nblobs=4, per_blob=5000, h=0.05, sigma_h=0.4, seed=3)
    rng = np.random.default_rng(8)
    h = 0.15
    pos = rng.uniform(-3 * h, 3 * h, size=(6000, 3)).astype(np.float32)
    vel = rng.normal(scale=0.3, size=(6000, 3)).astype(np.float32)
    small = sph_scenes.Scene("axis-cloud-6000", pos, vel, 0.02, 1000.0, 1.0, 1.04, h, -9.8, 0.2, 0.003,
                             (-8.0, 0.0, -8.0), (8.0, np.inf, 8.0))
    for scene in (big, small):
        print(f"noalloc vs allocating drop-in: {scene.name}, n={scene.n}", flush=True)
        ref, xf_ref, out, xf = _run_both(sph, scene)
        assert ref.tobytes() == out.tobytes(), f"{scene.name}: particles differ"
        assert xf_ref.tobytes() == xf.tobytes(), f"{scene.name}: transforms differ"
        assert np.all(np.isfinite(out["density"]))


def test_noalloc_rejects_sizes_beyond_the_maximum(sph):
    n = sph.NOALLOC_MAX_PARTICLES + 1
    p = sph.make_particles(np.zeros((n, 3), dtype=np.float32))
    with pytest.raises(sph.SphError) as err:
        sph.update_particles_noalloc(p, sph.settings(), 0.003, sph.box_reference_cpu())
    assert err.value.code == sph.SPH_E_LIMIT
