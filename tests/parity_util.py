"""Helpers shared by the GPU parity tests: run the oracle teacher-forced on a
state, map outputs by particle id, and the tolerances of the parity
contract.

Tolerances (FP32, per step, same input state on both sides):
  neighbour multisets     bit-exact
  density rho             |d rho| <= 1e-5 * rho_ref
  pressure p              |d p|   <= 1e-5 * k * max(rho_ref, rho0)
  force F                 |dF|    <= 1e-5 * sum_j m_ij (|f_press_ij| + |f_visc_ij|)
                          (interior forces are sums of ~30 nearly cancelling
                          terms; the bound is scaled by the term magnitudes)
  integrate / transforms  bit-exact when fed the same force and density
"""
from __future__ import annotations

import numpy as np

import sph_scenes

F32 = np.float32
RTOL = 1e-5


def oracle_step(orc, pos, vel, s, dt, box, nthreads=None):
    """Teacher-forced oracle step. Returns (by-id particle array, by-id
    transforms, neighbour CSR by id)."""
    n = len(pos)
    p = orc.particles(pos, vel, ids=np.arange(n))
    table = orc.prepare(p, s, nthreads)
    off, idx = orc.neighbors(p, table, s)
    orc.density(p, table, s, nthreads)
    orc.forces(p, table, s, nthreads)
    pre = p.copy()
    xf = orc.integrate(p, s, dt, box, nthreads)
    orc.free_table(table)
    ids = p["acceleration"][:, 0].astype(np.int64)
    out = np.zeros(n, dtype=p.dtype)
    out[ids] = p
    xfo = np.zeros_like(xf)
    xfo[ids] = xf
    rows = [None] * n
    for si in range(n):
        rows[ids[si]] = np.sort(ids[idx[off[si]:off[si + 1]]])
    return out, xfo, rows, pre, ids


def rows_from_csr(off, idx):
    return [idx[off[i]:off[i + 1]].astype(np.int64) for i in range(len(off) - 1)]


def force_scale(pos, vel, rho, prs, rows, s):
    """sum_j m_ij (|pressure term| + |viscosity term|) per particle, in
    double, from the oracle's own densities and neighbour multisets."""
    n = len(pos)
    scale = np.zeros(n)
    h = float(s.h)
    m = float(s.mass)
    mu = float(s.viscosity)
    sg = abs(float(s.spikyGrad))
    sl = float(s.spikyLap)
    lens = np.array([len(r) for r in rows])
    if lens.sum() == 0:
        return scale
    ii = np.repeat(np.arange(n), lens)
    jj = np.concatenate([r for r in rows if len(r)]).astype(np.int64)
    p64 = pos.astype(np.float64)
    d = p64[jj] - p64[ii]
    r = np.sqrt(np.sum(d * d, axis=1))
    hr = h - r
    fp = m * np.abs(prs[ii].astype(np.float64) + prs[jj]) / (2 * rho[jj]) * sg * hr * hr
    dv = np.linalg.norm(vel[jj].astype(np.float64) - vel[ii], axis=1)
    fv = mu * m * dv / rho[jj] * sl * hr
    np.add.at(scale, ii, fp + fv)
    return scale


def by_id(particles):
    """Reorders returned particles by the id carried in acceleration.x."""
    ids = particles["acceleration"][:, 0].astype(np.int64)
    out = np.zeros(len(particles), dtype=particles.dtype)
    out[ids] = particles
    return out, ids


def assert_rho_p(rho, prs, rho_ref, prs_ref, s, label=""):
    rho = np.asarray(rho, np.float64)
    rho_ref = np.asarray(rho_ref, np.float64)
    bad = np.abs(rho - rho_ref) > RTOL * np.abs(rho_ref)
    assert not bad.any(), f"{label} density: {bad.sum()} outside 1e-5, worst rel " \
                          f"{np.max(np.abs(rho - rho_ref) / np.abs(rho_ref)):.3e}"
    ptol = RTOL * float(s.gasConstant) * np.maximum(rho_ref, float(s.restDensity))
    badp = np.abs(np.asarray(prs, np.float64) - prs_ref) > ptol
    assert not badp.any(), f"{label} pressure: {badp.sum()} outside tolerance"


def assert_force(f, f_ref, scale, label=""):
    df = np.linalg.norm(np.asarray(f, np.float64) - f_ref, axis=1)
    tol = RTOL * scale + 1e-30
    bad = df > tol
    assert not bad.any(), f"{label} force: {bad.sum()} outside scaled tolerance, worst ratio " \
                          f"{np.max(df / np.maximum(tol, 1e-30)):.3e}"
    return float(np.max(df / np.maximum(tol, 1e-30))) if len(df) else 0.0


def collision_clump(n=600, h=0.05, seed=9):
    """A clump of n particles (side 0.4 h) around (0.3 h, h, -h): it spans
    reference cells (0,0,0), (0,1,-1), (0,1,0) and (0,0,-1).  The buckets of
    (0,-1,1) and (0,1,-1) coincide, so particles of cell (0,0,0) are slow
    and count their neighbours in (0,1,-1) twice.  Every lane's stencil
    holds the whole clump: waves whose lanes are all fast take the union
    walk, waves holding a slow lane the per-lane walk (which applies m_ij
    per neighbour), and n is small enough that no wave task is split."""
    rng = np.random.default_rng(seed)
    centre = np.array([0.3 * h, h, -h])
    pos = (centre + rng.uniform(-0.2 * h, 0.2 * h, size=(n, 3))).astype(F32)
    vel = rng.normal(0.0, 0.05, size=(n, 3)).astype(F32)
    mass = float(F32(0.02 * (h / 0.15) ** 3))
    return sph_scenes.Scene("collision-clump", pos, vel, mass, 1000.0, 1.0, 1.04, h, -9.8, 0.2,
                            0.001, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), steps=10)
