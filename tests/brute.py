"""Independent numpy restatements used to pin the oracle and the product's
host logic: brute-force O(N^2) neighbour multisets straight from the
definition of the reference scan, and closed forms."""
from __future__ import annotations

import math
from collections import Counter

import numpy as np

F32 = np.float32
HASH_MUL = (np.uint32(73856093), np.uint32(19349663), np.uint32(83492791))


def ref_cells(pos, h):
    """getCell: ivec3(pos / h) with float division and truncation toward
    zero (x86 semantics: NaN / out of range -> INT_MIN)."""
    q = np.asarray(pos, dtype=F32) / F32(h)
    out = np.full(q.shape, np.iinfo(np.int32).min, dtype=np.int64)
    ok = (q > F32(-2147483648.0)) & (q < F32(2147483648.0))
    out[ok] = np.trunc(q[ok]).astype(np.int64)
    return out


def h16(cells):
    """16-bit bucket of integer cells (n x 3)."""
    c = (np.asarray(cells, dtype=np.int64) & 0xFFFFFFFF).astype(np.uint32)
    with np.errstate(over="ignore"):
        v = (c[..., 0] * HASH_MUL[0]) ^ (c[..., 1] * HASH_MUL[1]) ^ (c[..., 2] * HASH_MUL[2])
    return (v & np.uint32(0xFFFF)).astype(np.uint32)


OFFSETS = np.array([(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1)], dtype=np.int64)


def dist2_rows(pos, i):
    """fl(fl(fl(dx*dx) + fl(dy*dy)) + fl(dz*dz)), dx = x_j - x_i, all j."""
    p = np.asarray(pos, dtype=F32)
    d = p - p[i]
    return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]


def neighbors_brute(pos, h):
    """Row i: indices j != i with fl(d^2) < fl(h*h), each repeated
    m_ij = #{o in {-1,0,1}^3 : H16(C_i + o) == H16(C_j)} times."""
    pos = np.asarray(pos, dtype=F32)
    h2 = F32(h) * F32(h)
    cells = ref_cells(pos, h)
    hj = h16(cells)
    rows = []
    for i in range(len(pos)):
        d2 = dist2_rows(pos, i)
        cand = np.nonzero(d2 < h2)[0]
        cand = cand[cand != i]
        around = Counter(h16(cells[i] + OFFSETS).tolist())
        row = []
        for j in cand:
            row.extend([int(j)] * around.get(int(hj[j]), 0))
        rows.append(np.array(sorted(row), dtype=np.int64))
    return rows


def buckets_collide_brute(cell):
    hs = h16(np.asarray(cell, dtype=np.int64) + OFFSETS)
    return len(set(hs.tolist())) < 27


def grid_coords(pos, origin, inv, dims):
    """Grid coordinate as the kernels compute it: floor((x - o) / H) in double,
    clamped to [0, dim), NaN -> 0."""
    q = (np.asarray(pos, dtype=np.float64) - np.asarray(origin, np.float64)) * inv
    out = np.zeros(q.shape, dtype=np.int64)
    for k in range(3):
        qk = q[:, k]
        g = np.where(qk >= 0, np.where(qk < dims[k], np.floor(np.where(np.isfinite(qk), qk, 0)), dims[k] - 1), 0)
        out[:, k] = g.astype(np.int64)
    return out


def poly6_w(r2, h):
    """m-free kernel value poly6 * (h^2 - r^2)^3 in double (closed form)."""
    return 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r2) ** 3


def reference_constants(mass, h):
    """SPHSettings constants in double with the reference's float PI."""
    pi = float(F32(3.14159265))
    poly6 = 315.0 / (64.0 * pi * h ** 9)
    spiky_grad = -45.0 / (pi * h ** 6)
    spiky_lap = 45.0 / (pi * h ** 6)
    return {"poly6": poly6, "spikyGrad": spiky_grad, "spikyLap": spiky_lap, "selfDens": mass * poly6 * h ** 6}


def forces_brute(pos, vel, rho, prs, rows, h, mass, visc, spiky_grad, spiky_lap, rho_of="j"):
    """Reference force (src/sph.cpp:107-121) in double from the definition:
    F_i = sum over the multiplicity-repeated row of i of
        -d m (p_i + p_j) / (2 rho_j) spikyGrad (h - r)^2
        + visc m (v_j - v_i) / rho_j spikyLap (h - r),
    d = (x_j - x_i) / r.  rho_of="i" divides by rho_i instead (a planted
    slip the pin must reject)."""
    pos = np.asarray(pos, np.float64)
    vel = np.asarray(vel, np.float64)
    rho = np.asarray(rho, np.float64)
    prs = np.asarray(prs, np.float64)
    out = np.zeros((len(pos), 3))
    for i, row in enumerate(rows):
        if len(row) == 0:
            continue
        j = np.asarray(row, np.int64)
        d = pos[j] - pos[i]
        r = np.sqrt(np.sum(d * d, axis=1))
        dirn = d / r[:, None]
        rd = rho[j] if rho_of == "j" else np.full(len(j), rho[i])
        fp = -dirn * (mass * (prs[i] + prs[j]) / (2.0 * rd) * spiky_grad * (h - r) ** 2)[:, None]
        fv = visc * mass * (vel[j] - vel[i]) / rd[:, None] * (spiky_lap * (h - r))[:, None]
        out[i] = np.sum(fp + fv, axis=0)
    return out
