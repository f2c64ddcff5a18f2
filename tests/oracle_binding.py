"""ctypes binding of the oracle (oracle/build/liborc.so): the CPU restatement
of the reference step, used only as the checker in tests and as the CPU
baseline in bench.py."""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_float, c_double, c_int, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(REPO, "oracle", "build", "liborc.so")


def default_threads() -> int:
    """Oracle worker threads: OMP_NUM_THREADS when set, else up to 16 cores."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(os.cpu_count() or 1, 16))

PARTICLE32 = np.dtype(
    {
        "names": ["position", "velocity", "acceleration", "force", "density", "pressure", "hash", "_pad"],
        "formats": [("<f4", 3), ("<f4", 3), ("<f4", 3), ("<f4", 3), "<f4", "<f4", "<u2", "<u2"],
        "offsets": [0, 12, 24, 36, 48, 52, 56, 58],
        "itemsize": 60,
    }
)
PARTICLE64 = np.dtype(
    {
        "names": ["position", "velocity", "acceleration", "force", "density", "pressure", "hash"],
        "formats": [("<f8", 3), ("<f8", 3), ("<f8", 3), ("<f8", 3), "<f8", "<f8", "<u2"],
        "offsets": [0, 24, 48, 72, 96, 104, 112],
        "itemsize": 120,
    }
)

_SETTING_NAMES = ("poly6", "spikyGrad", "spikyLap", "gasConstant", "mass", "h2", "selfDens",
                  "restDensity", "viscosity", "h", "g", "tension", "massPoly6Product")


class Settings32(ctypes.Structure):
    _fields_ = [("sphereScale", c_float * 16)] + [(n, c_float) for n in _SETTING_NAMES]


class Settings64(ctypes.Structure):
    _fields_ = [("sphereScale", c_double * 16)] + [(n, c_double) for n in _SETTING_NAMES]


class Box32(ctypes.Structure):
    _fields_ = [("lo", c_float * 3), ("hi", c_float * 3), ("restitution", c_float), ("offset", c_float)]


class Box64(ctypes.Structure):
    _fields_ = [("lo", c_double * 3), ("hi", c_double * 3), ("restitution", c_double), ("offset", c_double)]


class Oracle:
    """One flavour (32 or 64) of the oracle library."""

    def __init__(self, bits: int = 32, path: str = ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = ctypes.CDLL(path)
        self.bits = bits
        self.P = f"orc{bits}_"
        self.dtype = PARTICLE32 if bits == 32 else PARTICLE64
        self.Settings = Settings32 if bits == 32 else Settings64
        self.Box = Box32 if bits == 32 else Box64
        self.real = c_float if bits == 32 else c_double
        self.np_real = np.float32 if bits == 32 else np.float64
        R = self.real
        sig = {
            "settings_init": (None, [POINTER(self.Settings)] + [c_float] * 7),
            "box_reference_cpu": (None, [POINTER(self.Box)]),
            "get_hash": (c_uint32, [ctypes.c_int] * 3),
            "get_cell": (None, [POINTER(R), R, POINTER(ctypes.c_int)]),
            "calculate_hashes": (None, [c_void_p, c_size_t, POINTER(self.Settings), c_int]),
            "sort_particles": (None, [c_void_p, c_size_t]),
            "create_neighbor_table": (POINTER(c_uint32), [c_void_p, c_size_t]),
            "free_table": (None, [POINTER(c_uint32)]),
            "density_pressure": (None, [c_void_p, c_size_t, POINTER(c_uint32), POINTER(self.Settings), c_int]),
            "forces": (None, [c_void_p, c_size_t, POINTER(c_uint32), POINTER(self.Settings), c_int]),
            "integrate": (None, [c_void_p, c_size_t, c_void_p, POINTER(self.Settings), R, POINTER(self.Box), c_int]),
            "update_particles": (None, [c_void_p, c_void_p, c_size_t, POINTER(self.Settings), R, POINTER(self.Box),
                                        c_int]),
            "neighbors": (c_size_t, [c_void_p, c_size_t, POINTER(c_uint32), POINTER(self.Settings), c_void_p,
                                     c_void_p, c_size_t]),
            "density_subset": (None, [c_void_p, c_size_t, POINTER(c_uint32), POINTER(self.Settings), c_void_p,
                                      c_size_t, c_void_p, c_void_p]),
            "neighbors_subset": (c_size_t, [c_void_p, c_size_t, POINTER(c_uint32), POINTER(self.Settings),
                                            c_void_p, c_size_t, c_void_p, c_void_p, c_size_t]),
            "forces_subset": (None, [c_void_p, c_size_t, POINTER(c_uint32), POINTER(self.Settings), c_void_p,
                                     c_size_t, c_void_p]),
            "sizeof_particle": (c_int, []),
            "sizeof_settings": (c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(self.lib, self.P + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, "_" + name, fn)

    # ---- helpers --------------------------------------------------------
    def settings(self, mass, rest_density, gas_const, viscosity, h, g, tension):
        s = self.Settings()
        self._settings_init(ctypes.byref(s), mass, rest_density, gas_const, viscosity, h, g, tension)
        return s

    def box(self, lo, hi, restitution=0.5, offset=1e-4):
        b = self.Box()
        for k in range(3):
            b.lo[k] = lo[k]
            b.hi[k] = hi[k]
        b.restitution = restitution
        b.offset = offset
        return b

    def box_reference_cpu(self):
        b = self.Box()
        self._box_reference_cpu(ctypes.byref(b))
        return b

    def particles(self, pos, vel=None, ids=None):
        p = np.zeros(len(pos), dtype=self.dtype)
        p["position"] = np.asarray(pos, dtype=self.np_real)
        if vel is not None:
            p["velocity"] = np.asarray(vel, dtype=self.np_real)
        if ids is not None:  # ids ride in acceleration.x, untouched by the step
            p["acceleration"][:, 0] = np.asarray(ids, dtype=self.np_real)
        return p

    def get_hash(self, cx, cy, cz):
        return int(self._get_hash(cx, cy, cz))

    def get_cell(self, pos, h):
        arr = (self.real * 3)(*pos)
        out = (ctypes.c_int * 3)()
        self._get_cell(arr, h, out)
        return tuple(out)

    def update(self, p, s, dt, box, nthreads=None, transforms=True):
        """One full step in place on p; returns transforms (n x 16)."""
        nthreads = nthreads or default_threads()
        n = len(p)
        xf = np.zeros((n, 16), dtype=self.np_real) if transforms else None
        self._update_particles(p.ctypes.data_as(c_void_p), None if xf is None else xf.ctypes.data_as(c_void_p),
                               n, ctypes.byref(s), dt, ctypes.byref(box), nthreads)
        return xf

    def prepare(self, p, s, nthreads=None):
        """Hash, sort and build the table (in place); returns the table."""
        nthreads = nthreads or default_threads()
        n = len(p)
        self._calculate_hashes(p.ctypes.data_as(c_void_p), n, ctypes.byref(s), nthreads)
        self._sort_particles(p.ctypes.data_as(c_void_p), n)
        return self._create_neighbor_table(p.ctypes.data_as(c_void_p), n)

    def density(self, p, table, s, nthreads=None):
        nthreads = nthreads or default_threads()
        self._density_pressure(p.ctypes.data_as(c_void_p), len(p), table, ctypes.byref(s), nthreads)

    def forces(self, p, table, s, nthreads=None):
        nthreads = nthreads or default_threads()
        self._forces(p.ctypes.data_as(c_void_p), len(p), table, ctypes.byref(s), nthreads)

    def integrate(self, p, s, dt, box, nthreads=None, transforms=True):
        nthreads = nthreads or default_threads()
        n = len(p)
        xf = np.zeros((n, 16), dtype=self.np_real) if transforms else None
        self._integrate(p.ctypes.data_as(c_void_p), n, None if xf is None else xf.ctypes.data_as(c_void_p),
                        ctypes.byref(s), dt, ctypes.byref(box), nthreads)
        return xf

    def free_table(self, table):
        self._free_table(table)

    def neighbors(self, p_sorted, table, s):
        """CSR over sorted indices, each neighbour repeated m_ij times."""
        n = len(p_sorted)
        off = np.zeros(n + 1, dtype=np.uint64)
        total = self._neighbors(p_sorted.ctypes.data_as(c_void_p), n, table, ctypes.byref(s),
                                off.ctypes.data_as(c_void_p), None, 0)
        idx = np.zeros(max(total, 1), dtype=np.uint32)
        self._neighbors(p_sorted.ctypes.data_as(c_void_p), n, table, ctypes.byref(s),
                        off.ctypes.data_as(c_void_p), idx.ctypes.data_as(c_void_p), len(idx))
        return off, idx[:total]

    def density_subset(self, p_sorted, table, s, which):
        which = np.ascontiguousarray(which, dtype=np.uint32)
        rho = np.zeros(len(which), dtype=self.np_real)
        cnt = np.zeros(len(which), dtype=np.uint32)
        self._density_subset(p_sorted.ctypes.data_as(c_void_p), len(p_sorted), table, ctypes.byref(s),
                             which.ctypes.data_as(c_void_p), len(which), rho.ctypes.data_as(c_void_p),
                             cnt.ctypes.data_as(c_void_p))
        return rho, cnt

    def forces_subset(self, p_sorted, table, s, which):
        """Forces (k x 3) of the selected sorted indices, from the densities
        and pressures p_sorted holds."""
        which = np.ascontiguousarray(which, dtype=np.uint32)
        f = np.zeros((len(which), 3), dtype=self.np_real)
        self._forces_subset(p_sorted.ctypes.data_as(c_void_p), len(p_sorted), table, ctypes.byref(s),
                            which.ctypes.data_as(c_void_p), len(which), f.ctypes.data_as(c_void_p))
        return f

    def neighbors_subset(self, p_sorted, table, s, which):
        """CSR over the selected sorted indices: their rows of sorted
        indices, each neighbour repeated m_ij times."""
        which = np.ascontiguousarray(which, dtype=np.uint32)
        off = np.zeros(len(which) + 1, dtype=np.uint64)
        args = (p_sorted.ctypes.data_as(c_void_p), len(p_sorted), table, ctypes.byref(s),
                which.ctypes.data_as(c_void_p), len(which))
        total = self._neighbors_subset(*args, off.ctypes.data_as(c_void_p), None, 0)
        idx = np.zeros(max(total, 1), dtype=np.uint32)
        self._neighbors_subset(*args, off.ctypes.data_as(c_void_p), idx.ctypes.data_as(c_void_p), len(idx))
        return off, idx[:total]
