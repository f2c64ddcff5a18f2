"""ctypes binding of libsph_hip.so (the C ABI declared in include/sph_hip.h).

Mirrors the reference interface of lijenicol/SPH-Fluid-Simulator for the SPH
step:  ``settings(...)`` is ``SPHSettings(mass, restDensity, gasConst,
viscosity, h, g, tension)`` (src/SPHSystem.h:15-17), ``update_particles`` is
``updateParticles(particles, transforms, n, settings, dt, onGPU=true)``
(src/sph.h:19-22) with the CPU path's semantics, and ``Context`` is the
resident (state kept in HBM) form of ``SPHSystem::update``.

The library must be built (``make`` in the package directory); loading fails
loudly otherwise.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGE_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PACKAGE_DIR, "lib", "libsph_hip.so")

SPH_OK = 0
SPH_E_INVALID = -1
SPH_E_NODEVICE = -2
SPH_E_NOMEM = -3
SPH_E_HIP = -4
SPH_E_LIMIT = -5
SPH_E_STATE = -6
SPH_SCENE_AUTO_GRID = 4
SPH_SERVER_NO_BACKFILL = 2
SPH_SERVER_FAIL_RELAYOUT = 1
SPH_SERVER_FAIL_UPLOAD = 2
NOALLOC_MAX_PARTICLES = 65536  # SPH_NOALLOC_MAX_PARTICLES
SPH_SCENE_KEEP_FORCE = 1
SPH_SCENE_LANE_WALK = 2
SPH_SERVER_NO_GRAPH = 1
SPH_REQ_QUEUED, SPH_REQ_ACTIVE, SPH_REQ_DONE, SPH_REQ_FETCHED = 0, 1, 2, 3

# struct Particle (src/Particle.h:4-10): 60 bytes.
PARTICLE_DTYPE = np.dtype(
    {
        "names": ["position", "velocity", "acceleration", "force", "density", "pressure", "hash", "_pad"],
        "formats": [("<f4", 3), ("<f4", 3), ("<f4", 3), ("<f4", 3), "<f4", "<f4", "<u2", "<u2"],
        "offsets": [0, 12, 24, 36, 48, 52, 56, 58],
        "itemsize": 60,
    }
)


class Settings(ctypes.Structure):
    """struct SPHSettings (src/SPHSystem.h:13-22), 116 bytes."""

    _fields_ = [("sphereScale", c_float * 16)] + [
        (name, c_float)
        for name in (
            "poly6", "spikyGrad", "spikyLap", "gasConstant", "mass", "h2", "selfDens",
            "restDensity", "viscosity", "h", "g", "tension", "massPoly6Product",
        )
    ]


class Box(ctypes.Structure):
    _fields_ = [("lo", c_float * 3), ("hi", c_float * 3), ("restitution", c_float), ("offset", c_float)]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("n", c_uint32),
        ("dt", c_float),
        ("settings", Settings),
        ("box", Box),
        ("cell_scale", c_float),
        ("flags", c_uint32),
    ]


class Tune(ctypes.Structure):
    _fields_ = [("stencil_block", ctypes.c_int32), ("lane_walk", ctypes.c_int32), ("trials", ctypes.c_int32),
                ("cell_scale", ctypes.c_float), ("ms_per_step", c_double), ("default_ms_per_step", c_double),
                ("trial_ms", c_double * 6)]

    def as_dict(self):
        return {"stencil_block": self.stencil_block, "lane_walk": self.lane_walk, "trials": self.trials,
                "cell_scale": self.cell_scale, "ms_per_step": self.ms_per_step, "default_ms_per_step": self.default_ms_per_step,
                "trial_ms": list(self.trial_ms)}


class ServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget", c_uint64), ("max_scenes", c_uint32), ("max_steps_per_tick", c_uint32),
                ("flags", c_uint32), ("max_queued", c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, c_uint64) for k in ("submitted", "admitted", "completed", "queued", "active", "steps",
                                        "used_bytes", "budget_bytes")] + [
        (k, c_double) for k in ("particle_steps", "scene_steps", "busy_ms", "wall_s", "step_ms_p50",
                                "step_ms_p99", "step_ms_max")] + [(k, c_uint64) for k in ("rejected", "backfilled")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SphError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"sph_hip error {code}: {message}")
        self.code = code


_lib = None


def load(path: str | None = None) -> ctypes.CDLL:
    """Loads the HIP library (once) and declares the C signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is not built; run `make` in {PACKAGE_DIR}")
    lib = ctypes.CDLL(p)
    sig = {
        "sph_abi_version": (c_int, []),
        "sph_hip_last_error": (c_char_p, []),
        "sph_settings_init": (c_int, [POINTER(Settings)] + [c_float] * 7),
        "sph_box_reference_cpu": (c_int, [POINTER(Box)]),
        "sph_box_reference_gpu": (c_int, [POINTER(Box)]),
        "sph_get_hash": (c_uint32, [ctypes.c_int32] * 3),
        "sph_classify_host": (c_int, [POINTER(c_float), c_float, POINTER(ctypes.c_int32)]),
        "sph_hip_update_particles": (c_int, [c_void_p, c_void_p, c_size_t, POINTER(Settings), c_float]),
        "sph_hip_update_particles_box": (
            c_int, [c_void_p, c_void_p, c_size_t, POINTER(Settings), c_float, POINTER(Box)]),
        "sph_hip_update_particles_noalloc": (
            c_int, [c_void_p, c_void_p, c_size_t, POINTER(Settings), c_float, POINTER(Box)]),
        "sph_hip_neighbors": (
            c_int, [c_void_p, c_size_t, POINTER(Settings), c_void_p, c_void_p, c_size_t, POINTER(c_size_t)]),
        "sph_hip_create": (c_int, [POINTER(SceneDesc), POINTER(c_void_p)]),
        "sph_hip_destroy": (c_int, [c_void_p]),
        "sph_hip_upload": (c_int, [c_void_p, c_void_p, c_void_p]),
        "sph_hip_step": (c_int, [c_void_p, c_int]),
        "sph_hip_sync": (c_int, [c_void_p]),
        "sph_hip_download": (c_int, [c_void_p] * 6),
        "sph_hip_download_particles": (c_int, [c_void_p, c_void_p, c_void_p]),
        "sph_hip_set_graph": (c_int, [c_void_p, c_int]),
        "sph_hip_walk_dump": (c_int, [c_void_p] * 5),
        "sph_hip_grid_info": (c_int, [c_void_p, POINTER(c_float), POINTER(c_double)]),
        "sph_batch_download_particles": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p]),
        "sph_batch_grid_info": (c_int, [c_void_p, c_uint32, POINTER(c_float), POINTER(c_double)]),
        "sph_hip_unordered_cells": (c_int, [c_void_p, POINTER(c_uint64)]),
        "sph_batch_unordered_cells": (c_int, [c_void_p, POINTER(c_uint64)]),
        "sph_batch_walk_dump": (c_int, [c_void_p, c_uint32] + [c_void_p] * 4),
        "sph_hip_count": (c_uint32, [c_void_p]),
        "sph_hip_profile": (
            c_int, [c_void_p, c_int, POINTER(c_double), c_int, POINTER(c_int), c_char_p, c_size_t]),
        "sph_hip_copy_bench": (c_int, [c_size_t, c_int, POINTER(c_double)]),
        "sph_batch_create": (c_int, [POINTER(SceneDesc), c_uint32, POINTER(c_void_p)]),
        "sph_batch_destroy": (c_int, [c_void_p]),
        "sph_batch_upload": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p]),
        "sph_batch_step": (c_int, [c_void_p, c_int]),
        "sph_batch_sync": (c_int, [c_void_p]),
        "sph_batch_download": (c_int, [c_void_p, c_uint32] + [c_void_p] * 5),
        "sph_batch_set_graph": (c_int, [c_void_p, c_int]),
        "sph_batch_scene_count": (c_uint32, [c_void_p]),
        "sph_batch_profile": (
            c_int, [c_void_p, c_int, POINTER(c_double), c_int, POINTER(c_int), c_char_p, c_size_t]),
        "sph_batch_step_timed": (c_int, [c_void_p, c_int, c_void_p]),
        "sph_batch_autotune": (c_int, [c_void_p, c_int, POINTER(Tune)]),
        "sph_hip_autotune": (c_int, [c_void_p, c_int, POINTER(Tune)]),
        "sph_batch_autotune_grid": (c_int, [c_void_p, c_int, POINTER(Tune)]),
        "sph_hip_autotune_grid": (c_int, [c_void_p, c_int, POINTER(Tune)]),
        "sph_server_create": (c_int, [POINTER(ServerConfig), POINTER(c_void_p)]),
        "sph_server_destroy": (c_int, [c_void_p]),
        "sph_server_submit": (c_int, [c_void_p, POINTER(SceneDesc), c_void_p, c_void_p, c_uint32,
                                      POINTER(c_uint64)]),
        "sph_server_tick": (c_int, [c_void_p, POINTER(c_uint32)]),
        "sph_server_status": (c_int, [c_void_p, c_uint64, POINTER(c_int), POINTER(c_uint32),
                                      POINTER(c_double)]),
        "sph_server_fetch": (c_int, [c_void_p, c_uint64] + [c_void_p] * 5),
        "sph_server_inject_failure": (c_int, [c_void_p, c_int]),
        "sph_server_stats": (c_int, [c_void_p, POINTER(ServerStats)]),
        "sph_server_step_times": (c_int, [c_void_p, c_void_p, c_size_t, POINTER(c_size_t)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != SPH_OK:
        msg = load().sph_hip_last_error()
        raise SphError(rc, msg.decode() if msg else "")


def _ptr(a: np.ndarray | None) -> c_void_p | None:
    return None if a is None else a.ctypes.data_as(c_void_p)


def settings(mass=0.02, rest_density=1000.0, gas_const=1.0, viscosity=1.04, h=0.15, g=-9.8,
             tension=0.2) -> Settings:
    """SPHSettings(mass, restDensity, gasConst, viscosity, h, g, tension)."""
    s = Settings()
    _check(load().sph_settings_init(ctypes.byref(s), mass, rest_density, gas_const, viscosity, h, g, tension))
    return s


def box_reference_cpu() -> Box:
    b = Box()
    _check(load().sph_box_reference_cpu(ctypes.byref(b)))
    return b


def box_reference_gpu() -> Box:
    b = Box()
    _check(load().sph_box_reference_gpu(ctypes.byref(b)))
    return b


def box(lo, hi, restitution=0.5, offset=1e-4) -> Box:
    b = Box()
    for k in range(3):
        b.lo[k] = lo[k]
        b.hi[k] = hi[k]
    b.restitution = restitution
    b.offset = offset
    return b


def get_hash(cx: int, cy: int, cz: int) -> int:
    return int(load().sph_get_hash(cx, cy, cz))


def classify_host(pos, h: float) -> tuple[int, bool, tuple[int, int, int]]:
    p = (c_float * 3)(*[float(v) for v in pos])
    out = (ctypes.c_int32 * 5)()
    _check(load().sph_classify_host(p, h, out))
    return int(out[0]), bool(out[1]), (int(out[2]), int(out[3]), int(out[4]))


def make_particles(pos: np.ndarray, vel: np.ndarray | None = None) -> np.ndarray:
    """Reference Particle records from n x 3 positions (and velocities)."""
    pos = np.asarray(pos, dtype=np.float32)
    p = np.zeros(len(pos), dtype=PARTICLE_DTYPE)
    p["position"] = pos
    if vel is not None:
        p["velocity"] = np.asarray(vel, dtype=np.float32)
    return p


def update_particles(particles: np.ndarray, s: Settings, dt: float, box_: Box | None = None,
                     transforms: bool = True):
    """One drop-in step, in place on a copy.  Returns (particles, transforms)."""
    p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE).copy()
    n = len(p)
    xf = np.zeros((n, 16), dtype=np.float32) if transforms else None
    lib = load()
    if box_ is None:
        rc = lib.sph_hip_update_particles(_ptr(p), _ptr(xf), n, ctypes.byref(s), dt)
    else:
        rc = lib.sph_hip_update_particles_box(_ptr(p), _ptr(xf), n, ctypes.byref(s), dt, ctypes.byref(box_))
    _check(rc)
    return p, xf


def update_particles_noalloc(particles: np.ndarray, s: Settings, dt: float, box_: Box,
                             transforms: bool = True):
    """As update_particles with a box, through the entry point that
    allocates no device memory (n <= NOALLOC_MAX_PARTICLES)."""
    p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE).copy()
    n = len(p)
    xf = np.zeros((n, 16), dtype=np.float32) if transforms else None
    _check(load().sph_hip_update_particles_noalloc(_ptr(p), _ptr(xf), n, ctypes.byref(s), dt,
                                                   ctypes.byref(box_)))
    return p, xf


def neighbors(particles: np.ndarray, s: Settings):
    """Neighbour multisets (CSR over input indices, multiplicity-repeated)."""
    p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
    n = len(p)
    off = np.zeros(n + 1, dtype=np.uint64)
    total = c_size_t(0)
    lib = load()
    _check(lib.sph_hip_neighbors(_ptr(p), n, ctypes.byref(s), _ptr(off), None, 0, ctypes.byref(total)))
    idx = np.zeros(max(int(total.value), 1), dtype=np.uint32)
    _check(lib.sph_hip_neighbors(_ptr(p), n, ctypes.byref(s), _ptr(off), _ptr(idx), len(idx),
                                 ctypes.byref(total)))
    return off, idx[: int(total.value)]


class Context:
    """Resident single scene: SPHSystem::update without per-step PCIe."""

    def __init__(self, n: int, s: Settings, box_: Box, dt: float, cell_scale: float = 1.0,
                 keep_force: bool = False, lane_walk: bool = False, auto_grid: bool = False):
        self._lib = load()
        d = scene_desc(n, s, box_, dt, cell_scale, keep_force, lane_walk, auto_grid)
        h = c_void_p()
        _check(self._lib.sph_hip_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.n = n

    def close(self):
        if self._h:
            self._lib.sph_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, pos: np.ndarray, vel: np.ndarray | None = None):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(self.n, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(self.n, 3)
        _check(self._lib.sph_hip_upload(self._h, _ptr(pos), _ptr(vel)))

    def step(self, nsteps: int = 1):
        _check(self._lib.sph_hip_step(self._h, nsteps))

    def sync(self):
        _check(self._lib.sph_hip_sync(self._h))

    def set_graph(self, enable: bool):
        _check(self._lib.sph_hip_set_graph(self._h, 1 if enable else 0))

    def download(self):
        n = self.n
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        prs = np.zeros(n, np.float32)
        ids = np.zeros(n, np.uint32)
        _check(self._lib.sph_hip_download(self._h, _ptr(pos), _ptr(vel), _ptr(rho), _ptr(prs), _ptr(ids)))
        return {"pos": pos, "vel": vel, "rho": rho, "pressure": prs, "ids": ids}

    def download_particles(self, transforms: bool = True):
        p = np.zeros(self.n, dtype=PARTICLE_DTYPE)
        xf = np.zeros((self.n, 16), np.float32) if transforms else None
        _check(self._lib.sph_hip_download_particles(self._h, _ptr(p), _ptr(xf)))
        return p, xf

    def autotune(self, trial_steps: int = 10, grid: bool = False) -> dict:
        """Picks the stencil block size (bit-preserving); with grid=True also
        the cell scale, 1 or 1/2 (changes results by rounding)."""
        t = Tune()
        fn = self._lib.sph_hip_autotune_grid if grid else self._lib.sph_hip_autotune
        _check(fn(self._h, trial_steps, ctypes.byref(t)))
        return t.as_dict()

    def grid_info(self):
        """(current cell scale, candidates per neighbour of the last step's
        state on the current grid)."""
        sc, r = c_float(0), c_double(0)
        _check(self._lib.sph_hip_grid_info(self._h, ctypes.byref(sc), ctypes.byref(r)))
        return float(sc.value), float(r.value)

    def unordered_cells(self) -> int:
        """Cells of more than 16,384 particles that kept arrival order (not
        bit-reproducible), summed over every cell sort so far."""
        v = c_uint64(0)
        _check(self._lib.sph_hip_unordered_cells(self._h, ctypes.byref(v)))
        return int(v.value)

    def walk_dump(self, lists: bool = False) -> dict:
        """Neighbour counts and walk classes of the production walks on the
        current state (sph_hip_walk_dump), by particle id."""
        return _walk_dump(lambda *a: self._lib.sph_hip_walk_dump(self._h, *a), self.n, lists)

    def profile(self, nsteps: int = 10) -> dict:
        ms = (c_double * 32)()
        nk = c_int(0)
        names = ctypes.create_string_buffer(1024)
        _check(self._lib.sph_hip_profile(self._h, nsteps, ms, 32, ctypes.byref(nk), names, 1024))
        labels = names.raw.split(b"\0")[: nk.value]
        return {labels[k].decode(): float(ms[k]) for k in range(nk.value)}


LIST_MAX = 64  # neighbour-list entries per particle (kListMax)


def _walk_dump(fn, n: int, lists: bool):
    count = np.zeros(n, np.uint32)
    walk = np.zeros(n, np.uint32)
    lcount = np.zeros(n, np.uint32)
    lst = np.zeros((n, LIST_MAX), np.uint32) if lists else None
    _check(fn(_ptr(count), _ptr(walk), _ptr(lcount), _ptr(lst)))
    return {"count": count, "union": (walk & 1) != 0, "split": (walk & 2) != 0, "listed": (walk & 4) != 0,
            "list_count": lcount, "list": lst}


def scene_desc(n: int, s: Settings, box_: Box, dt: float, cell_scale: float = 1.0,
               keep_force: bool = False, lane_walk: bool = False, auto_grid: bool = False) -> SceneDesc:
    d = SceneDesc()
    d.n = n
    d.dt = dt
    d.settings = s
    d.box = box_
    d.cell_scale = cell_scale
    d.flags = ((SPH_SCENE_KEEP_FORCE if keep_force else 0) | (SPH_SCENE_LANE_WALK if lane_walk else 0) |
               (SPH_SCENE_AUTO_GRID if auto_grid else 0))
    return d


class Batch:
    """Independent scenes advanced together, one launch per pass."""

    def __init__(self, descs):
        self._lib = load()
        arr = (SceneDesc * len(descs))(*descs)
        h = c_void_p()
        _check(self._lib.sph_batch_create(arr, len(descs), ctypes.byref(h)))
        self._h = h
        self.counts = [d.n for d in descs]

    def close(self):
        if self._h:
            self._lib.sph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, scene: int, pos: np.ndarray, vel: np.ndarray | None = None):
        n = self.counts[scene]
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(n, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        _check(self._lib.sph_batch_upload(self._h, scene, _ptr(pos), _ptr(vel)))

    def step(self, nsteps: int = 1):
        _check(self._lib.sph_batch_step(self._h, nsteps))

    def sync(self):
        _check(self._lib.sph_batch_sync(self._h))

    def set_graph(self, enable: bool):
        _check(self._lib.sph_batch_set_graph(self._h, 1 if enable else 0))

    def download(self, scene: int):
        n = self.counts[scene]
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        prs = np.zeros(n, np.float32)
        ids = np.zeros(n, np.uint32)
        _check(self._lib.sph_batch_download(self._h, scene, _ptr(pos), _ptr(vel), _ptr(rho), _ptr(prs),
                                            _ptr(ids)))
        return {"pos": pos, "vel": vel, "rho": rho, "pressure": prs, "ids": ids}

    def step_timed(self, nsteps: int) -> np.ndarray:
        ms = np.zeros(nsteps, np.float32)
        _check(self._lib.sph_batch_step_timed(self._h, nsteps, _ptr(ms)))
        return ms

    def autotune(self, trial_steps: int = 10, grid: bool = False) -> dict:
        """As Context.autotune, for every scene of the batch together."""
        t = Tune()
        fn = self._lib.sph_batch_autotune_grid if grid else self._lib.sph_batch_autotune
        _check(fn(self._h, trial_steps, ctypes.byref(t)))
        return t.as_dict()

    def download_particles(self, scene: int, transforms: bool = True):
        """Reference-layout particles of one scene (device order) and, if
        asked, its transforms."""
        n = self.counts[scene]
        p = np.zeros(n, dtype=PARTICLE_DTYPE)
        xf = np.zeros((n, 16), dtype=np.float32) if transforms else None
        _check(self._lib.sph_batch_download_particles(self._h, scene, _ptr(p), _ptr(xf)))
        return p, xf

    def grid_info(self, scene: int):
        """Context.grid_info for one scene."""
        sc, r = c_float(0), c_double(0)
        _check(self._lib.sph_batch_grid_info(self._h, scene, ctypes.byref(sc), ctypes.byref(r)))
        return float(sc.value), float(r.value)

    def unordered_cells(self) -> int:
        """Context.unordered_cells over all scenes of the batch."""
        v = c_uint64(0)
        _check(self._lib.sph_batch_unordered_cells(self._h, ctypes.byref(v)))
        return int(v.value)

    def walk_dump(self, scene: int, lists: bool = False) -> dict:
        """As Context.walk_dump, for one scene of the batch."""
        return _walk_dump(lambda *a: self._lib.sph_batch_walk_dump(self._h, scene, *a), self.counts[scene], lists)

    def profile(self, nsteps: int = 10) -> dict:
        ms = (c_double * 32)()
        nk = c_int(0)
        names = ctypes.create_string_buffer(1024)
        _check(self._lib.sph_batch_profile(self._h, nsteps, ms, 32, ctypes.byref(nk), names, 1024))
        labels = names.raw.split(b"\0")[: nk.value]
        return {labels[k].decode(): float(ms[k]) for k in range(nk.value)}


class Server:
    """Continuous-batching serving loop (sph_server_*): submit scene-advance
    requests, call tick() to run scheduling rounds, fetch finished results."""

    def __init__(self, memory_budget: int = 0, max_scenes: int = 0, max_steps_per_tick: int = 0,
                 graph: bool = True, max_queued: int = 0, backfill: bool = True):
        self._lib = load()
        flags = (0 if graph else SPH_SERVER_NO_GRAPH) | (0 if backfill else SPH_SERVER_NO_BACKFILL)
        cfg = ServerConfig(memory_budget, max_scenes, max_steps_per_tick, flags, max_queued)
        h = c_void_p()
        _check(self._lib.sph_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.counts = {}

    def close(self):
        if self._h:
            self._lib.sph_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, desc: SceneDesc, pos: np.ndarray, vel: np.ndarray | None, steps: int) -> int:
        n = desc.n
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(n, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        rid = c_uint64(0)
        _check(self._lib.sph_server_submit(self._h, ctypes.byref(desc), _ptr(pos), _ptr(vel), steps,
                                           ctypes.byref(rid)))
        self.counts[rid.value] = n
        return rid.value

    def tick(self) -> int:
        ran = c_uint32(0)
        _check(self._lib.sph_server_tick(self._h, ctypes.byref(ran)))
        return ran.value

    def status(self, rid: int):
        st = c_int(0)
        done = c_uint32(0)
        t = (c_double * 3)()
        _check(self._lib.sph_server_status(self._h, rid, ctypes.byref(st), ctypes.byref(done), t))
        return st.value, done.value, (t[0], t[1], t[2])

    def fetch(self, rid: int):
        n = self.counts[rid]
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        rho = np.zeros(n, np.float32)
        prs = np.zeros(n, np.float32)
        ids = np.zeros(n, np.uint32)
        _check(self._lib.sph_server_fetch(self._h, rid, _ptr(pos), _ptr(vel), _ptr(rho), _ptr(prs), _ptr(ids)))
        return {"pos": pos, "vel": vel, "rho": rho, "pressure": prs, "ids": ids}

    def discard(self, rid: int):
        """Releases a finished request's result without copying it."""
        _check(self._lib.sph_server_fetch(self._h, rid, None, None, None, None, None))

    def stats(self) -> dict:
        st = ServerStats()
        _check(self._lib.sph_server_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def inject_failure(self, stage: int):
        """Tests only: the next relayout (1) or upload (2) fails once."""
        _check(self._lib.sph_server_inject_failure(self._h, stage))

    def step_times(self) -> np.ndarray:
        n = c_size_t(0)
        _check(self._lib.sph_server_step_times(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, np.float32)
        _check(self._lib.sph_server_step_times(self._h, _ptr(out), n.value, ctypes.byref(n)))
        return out


def copy_bench(nbytes: int, iters: int = 20) -> float:
    """Mean milliseconds of one device copy of nbytes (read nbytes + write nbytes)."""
    ms = c_double(0.0)
    _check(load().sph_hip_copy_bench(nbytes, iters, ctypes.byref(ms)))
    return ms.value
