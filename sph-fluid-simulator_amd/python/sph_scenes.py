"""Scene generators for the SPH step: the reference's own initial state and
the benchmark workloads (jittered lattice blocks).

``reference_init`` reproduces SPHSystem::initParticles
(src/SPHSystem.cpp:76-108 of lijenicol/SPH-Fluid-Simulator): srand(1024),
separation h + 0.01, one-sided jitter in [-h/10, -h/20], origin
(-1.5, h + 0.1, -1.5), index i + (j + w k) w.  It calls the C library's
srand/rand, so it matches the reference on the same libc.

The numbered workloads are:
  0  8,000 particles (20^3, spacing 0.5 h, jitter +-0.01 h), h = 0.15,
     cube centred at (0, 0.5, 0) in the box [-1,1]^3 without ceiling,
     dt = 0.003, 1000 steps
  1  262,144 particles (64^3 dam break), h = 0.05, mass 0.02 (h / 0.15)^3,
     block anchored at the corner (-1, -1, -1) of [-1,1]^3, dt = 0.003
  2  1,000,000 particles (100^3 dam break), h = 0.03, mass 1.6e-4,
     box [-2,2] x [-1,1]^2 anchored at its low corner, dt = 0.001
Lattice blocks anchored at a corner start half a spacing inside the
reflection planes lo + h, so no particle starts beyond a wall.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


@dataclass
class Scene:
    name: str
    pos: np.ndarray  # n x 3 float32
    vel: np.ndarray  # n x 3 float32
    mass: float
    rest_density: float
    gas_const: float
    viscosity: float
    h: float
    g: float
    tension: float
    dt: float
    box_lo: tuple
    box_hi: tuple
    restitution: float = 0.5
    offset: float = 1e-4
    steps: int = 1000
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.pos)

    def settings_args(self):
        return (self.mass, self.rest_density, self.gas_const, self.viscosity, self.h, self.g, self.tension)


def lattice(counts, spacing, origin, jitter, seed):
    """Jittered lattice, x fastest: index a + nx (b + ny c).  Jitter is
    uniform in [-jitter, jitter) per axis from numpy's PCG64(seed)."""
    nx, ny, nz = counts
    a = np.arange(nx, dtype=np.float64)
    b = np.arange(ny, dtype=np.float64)
    c = np.arange(nz, dtype=np.float64)
    C, B, A = np.meshgrid(c, b, a, indexing="ij")
    base = np.stack([A.ravel(), B.ravel(), C.ravel()], axis=1) * spacing + np.asarray(origin, np.float64)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=base.shape)
    return (base + jit).astype(F32)


def reference_init(cube_width: int, h: float):
    """SPHSystem::initParticles, src/SPHSystem.cpp:76-108 (libc rand)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(1024)
    rand_max = F32(2147483647)
    h = F32(h)
    sep = h + F32(0.01)
    w = cube_width
    pos = np.zeros((w * w * w, 3), dtype=F32)
    for i in range(w):
        for j in range(w):
            for k in range(w):
                r = []
                for _ in range(3):
                    u = F32(libc.rand()) / rand_max
                    r.append((u * F32(0.5) - F32(1)) * h / F32(10))
                x = F32(i) * sep + r[0] - F32(1.5)
                y = F32(j) * sep + r[1] + h + F32(0.1)
                z = F32(k) * sep + r[2] - F32(1.5)
                pos[i + (j + w * k) * w] = (x, y, z)
    return pos


def workload(k: int, seed: int = 1024) -> Scene:
    if k == 0:
        h, sp, n1 = 0.15, 0.075, 20
        origin = [0.0 - 0.5 * (n1 - 1) * sp, 0.5 - 0.5 * (n1 - 1) * sp, 0.0 - 0.5 * (n1 - 1) * sp]
        pos = lattice((n1, n1, n1), sp, origin, 0.01 * h, seed)
        return Scene("workload0-cube-8k", pos, np.zeros_like(pos), 0.02, 1000.0, 1.0, 1.04, h, -9.8, 0.2,
                     0.003, (-1.0, -1.0, -1.0), (1.0, math.inf, 1.0), steps=1000)
    if k == 1:
        h, sp, n1 = 0.05, 0.025, 64
        o = -1.0 + h + 0.5 * sp
        pos = lattice((n1, n1, n1), sp, [o, o, o], 0.01 * h, seed)
        mass = float(F32(0.02 * (h / 0.15) ** 3))
        return Scene("workload1-dambreak-262k", pos, np.zeros_like(pos), mass, 1000.0, 1.0, 1.04, h, -9.8, 0.2,
                     0.003, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), steps=2000)
    if k == 2:
        h, sp, n1 = 0.03, 0.015, 100
        o = [-2.0 + h + 0.5 * sp, -1.0 + h + 0.5 * sp, -1.0 + h + 0.5 * sp]
        pos = lattice((n1, n1, n1), sp, o, 0.01 * h, seed)
        return Scene("workload2-dambreak-1M", pos, np.zeros_like(pos), 1.6e-4, 1000.0, 1.0, 1.04, h, -9.8, 0.2,
                     0.001, (-2.0, -1.0, -1.0), (2.0, 1.0, 1.0), steps=1000)
    raise ValueError(f"unknown workload {k}")


def reference_app_scene(cube_width: int = 15) -> Scene:
    """The reference application's default system: SPHSettings(0.02, 1000,
    1, 1.04, 0.15, -9.8, 0.2) and SPHSystem(15, ...) (src/Tester.cpp:90-91),
    in the reference CPU box, dt = 0.003 (src/SPHSystem.cpp:113)."""
    h = 0.15
    pos = reference_init(cube_width, h)
    return Scene(f"reference-app-{cube_width}^3", pos, np.zeros_like(pos), 0.02, 1000.0, 1.0, 1.04, h, -9.8, 0.2,
                 0.003, (-8.0, 0.0, -8.0), (8.0, math.inf, 8.0), steps=1000)


def batched_workload(nscenes: int = 256, seed: int = 1024, dt: float = 0.002):
    """The batched workload: `nscenes` independent scenes, particle counts
    log-uniform in [1k, 64k], each a jittered lattice block (spacing
    0.8 / ceil(cbrt(N)), h = 2.1 spacing, i.e. ~32 interior neighbours) at a
    random corner of its own [-1,1]^3 box, viscosity ~ U[0.5, 2], gas
    constant ~ U[0.5, 3], mass 0.02 (h / 0.15)^3, 500 steps."""
    rng = np.random.default_rng(seed)
    scenes = []
    for k in range(nscenes):
        n = int(round(math.exp(rng.uniform(math.log(1000.0), math.log(65536.0)))))
        side = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
        sp = 0.8 / side
        h = 2.1 * sp
        corner = rng.integers(0, 2, size=3)
        lo = -1.0 + h + 0.5 * sp
        hi = 1.0 - h - 0.5 * sp - (side - 1) * sp
        origin = [lo if c == 0 else hi for c in corner]
        pos = lattice((side, side, side), sp, origin, 0.01 * h, seed + 7919 * (k + 1))[:n]
        mu = float(rng.uniform(0.5, 2.0))
        gas = float(rng.uniform(0.5, 3.0))
        mass = float(F32(0.02 * (h / 0.15) ** 3))
        scenes.append(Scene(f"batched-{k}-{n}", pos, np.zeros_like(pos), mass, 1000.0, gas, mu, h, -9.8, 0.2,
                            dt, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), steps=500))
    return scenes


def blobs(nblobs: int = 12, per_blob: int = 2000, h: float = 0.05, sigma_h: float = 0.5, seed: int = 31) -> Scene:
    """Gaussian clumps (sigma = sigma_h h) of per_blob particles each, at
    random centres in [-0.6, 0.6]^3 inside the box [-1,1]^3: the clumped
    state the dam-break workloads reach, with cells holding hundreds of
    particles and waves that span one to a few cells."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-0.6, 0.6, size=(nblobs, 3))
    pos = (np.repeat(centres, per_blob, axis=0) + rng.normal(0.0, sigma_h * h, size=(nblobs * per_blob, 3)))
    pos = np.clip(pos, -0.9, 0.9).astype(F32)
    vel = rng.normal(0.0, 0.05, size=pos.shape).astype(F32)
    mass = float(F32(0.02 * (h / 0.15) ** 3))
    return Scene(f"blobs-{nblobs}x{per_blob}", pos, vel, mass, 1000.0, 1.0, 1.04, h, -9.8, 0.2,
                 0.001, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), steps=200)


def request_scene(n: int, kind: str, seed: int, dt: float = 0.001) -> Scene:
    """One serving request's initial state in the box [-1,1]^3, h = 2.1
    spacing (~32 interior neighbours), mass 0.02 (h / 0.15)^3:
      dambreak  jittered lattice block (spacing 0.8 / ceil(cbrt n)) at the
                low corner, half a spacing inside the reflection planes;
      droplet   jittered lattice inside a sphere of radius 0.3 centred at
                (0, 0.6, 0), spacing chosen so the sphere holds at least n
                points, n of them kept at random."""
    if kind == "dambreak":
        side = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
        sp = 0.8 / side
        h = 2.1 * sp
        o = -1.0 + h + 0.5 * sp
        pos = lattice((side, side, side), sp, [o, o, o], 0.01 * h, seed)[:n]
    elif kind == "droplet":
        r = 0.3
        sp = (4.0 / 3.0 * math.pi * r ** 3 / n) ** (1.0 / 3.0)
        while True:
            m = int(math.ceil(2 * r / sp)) + 1
            g = lattice((m, m, m), sp, [-0.5 * (m - 1) * sp] * 3, 0.0, seed)
            inside = g[np.einsum("ij,ij->i", g.astype(np.float64), g.astype(np.float64)) <= r * r]
            if len(inside) >= n:
                break
            sp *= 0.99
        h = 2.1 * sp
        rng = np.random.default_rng(seed)
        keep = np.sort(rng.choice(len(inside), size=n, replace=False))  # whole sphere, lattice order
        pos = (inside[keep].astype(np.float64) + rng.uniform(-0.01 * h, 0.01 * h, size=(n, 3))
               + np.array([0.0, 0.6, 0.0])).astype(F32)
    else:
        raise ValueError(f"unknown request geometry {kind!r}")
    mass = float(F32(0.02 * (h / 0.15) ** 3))
    return Scene(f"{kind}-{n}", pos, np.zeros_like(pos), mass, 1000.0, 1.0, 1.04, h, -9.8, 0.2,
                 dt, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))


def request_stream(rate: float, duration: float, seed: int = 1024, n_lo: int = 2000, n_hi: int = 500000,
                   steps_lo: int = 10, steps_hi: int = 1000, max_requests: int | None = None):
    """Poisson arrivals over [0, duration) at `rate` per second (at most
    max_requests of them): a list of (arrival seconds, particle count
    log-uniform in [n_lo, n_hi], steps uniform in [steps_lo, steps_hi],
    geometry dambreak or droplet, seed).  Every field but the arrival time
    is drawn in the same order at any rate, so request i is the same
    request at every rate."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    while max_requests is None or len(out) < max_requests:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        kind = "dambreak" if rng.random() < 0.5 else "droplet"
        out.append((t, n, steps, kind, int(rng.integers(0, 2**31 - 1))))
    return out
