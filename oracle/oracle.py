"""ctypes wrapper of liboracle.so, the FP64 C restatement of the reference.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product never imports it.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ORACLE_DIR)
sys.path.insert(0, os.path.join(REPO, "multi-block-lbm-gpu_amd"))

from mbl.abi import AREA, Inf_Ff_store, Int3, mbl_config, mbl_geometry  # noqa: E402

LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")
_lib = None

FIELDS = {"solid": 0, "x": 1, "y": 2, "rho": 3, "ux": 4, "uy": 5, "forcex": 6, "forcey": 7,
          "upx": 26, "upy": 27}


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    sig = {
        "orc_feq": (C.c_double, [ip, C.c_double, C.c_double, C.c_double]),
        "orc_macros": (None, [dp, dp, dp, dp]),
        "orc_geo_init": (ip, [C.POINTER(mbl_geometry), vp, vp, vp, vp, vp]),
        "orc_sim_create": (vp, [C.POINTER(mbl_geometry), C.POINTER(mbl_config)]),
        "orc_sim_destroy": (None, [vp]),
        "orc_sim_run": (ip, [vp, ip]),
        "orc_sim_evolve": (ip, [vp, ip]),
        "orc_sim_diag": (ip, [vp, dp]),
        "orc_sim_poiseuille_l2": (C.c_double, [vp]),
        "orc_temporal_value": (C.c_double, [C.c_double] * 4),
        "orc_sim_steps_done": (ip, [vp]),
        "orc_sim_num_areas": (ip, [vp]),
        "orc_sim_num_lines": (ip, [vp, ip]),
        "orc_sim_area": (None, [vp, ip, C.POINTER(AREA)]),
        "orc_sim_line": (None, [vp, ip, ip, C.POINTER(Inf_Ff_store)]),
        "orc_sim_get": (ip, [vp, ip, ip, dp]),
        "orc_sim_set": (ip, [vp, ip, ip, dp]),
        "orc_sim_drag_links": (ip, [vp, C.POINTER(Int3), ip]),
        "orc_sim_series": (ip, [vp, ip, dp, ip]),
        "orc_sim_ring": (ip, [vp, ip, ip, ip, dp]),
        "orc_sim_markers": (None, [vp, ip, C.POINTER(C.c_int)]),
        "orc_sim_spline_M": (ip, [vp, ip, dp]),
        "orc_gtsv": (None, [ip, ip, dp, dp, dp, dp, ip]),
        "orc_tri_A": (None, [ip, dp, dp, dp]),
        "orc_spline3_s": (C.c_double, [C.c_double] * 7),
        "orc_spline_column": (None, [ip, ip, dp, dp, dp]),
        "orc_random_velocity": (None, [C.c_ulonglong, C.c_double, ip, ip, dp, dp]),
        "orc_periodic_create": (vp, [ip, ip, C.c_double, C.c_double, C.c_ulonglong, C.c_double]),
        "orc_periodic_destroy": (None, [vp]),
        "orc_periodic_run": (C.c_double, [vp, ip]),
        "orc_periodic_get": (ip, [vp, ip, dp]),
        "orc_threads": (ip, []),
        "orc_sim_set_block_hook": (None, [vp, vp, vp]),
        "orc_sim_spline_line": (None, [vp, ip]),
        "orc_sim_transfer_line": (None, [vp, ip, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


BLOCK_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def field_id(field):
    if isinstance(field, tuple):
        which, k = field
        return (8 if which == "f" else 17) + k
    return FIELDS[field]


class OracleSim:
    """Whole-run restatement of Multi-block-LBM.cu:main on the CPU in FP64."""

    def __init__(self, geometry, cfg):
        self.lib = load()
        self._g = geometry  # keep the tables alive
        self.p = self.lib.orc_sim_create(C.byref(geometry), C.byref(cfg))
        self.na = self.lib.orc_sim_num_areas(self.p)
        self.areas = []
        for a in range(self.na):
            ar = AREA()
            self.lib.orc_sim_area(self.p, a, C.byref(ar))
            self.areas.append(ar)

    def run(self, n):
        self.lib.orc_sim_run(self.p, n)

    def evolve(self, n):
        self.lib.orc_sim_evolve(self.p, n)

    def diag(self):
        out = np.zeros(3)
        self.lib.orc_sim_diag(self.p, _dptr(out))
        return out

    def poiseuille_l2(self):
        return self.lib.orc_sim_poiseuille_l2(self.p)

    def shape(self, a):
        return self.areas[a].ny, self.areas[a].nx

    def get(self, a, field):
        ny, nx = self.shape(a)
        out = np.zeros(nx * ny)
        rc = self.lib.orc_sim_get(self.p, a, field_id(field), _dptr(out))
        assert rc == 0
        return out.reshape(ny, nx)

    def set(self, a, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        assert self.lib.orc_sim_set(self.p, a, field_id(field), _dptr(v)) == 0

    def series(self, kind):
        n = self.lib.orc_sim_series(self.p, kind, None, 0)
        out = np.zeros(max(1, n))
        self.lib.orc_sim_series(self.p, kind, _dptr(out), n)
        return out[:n]

    def links(self):
        n = self.lib.orc_sim_drag_links(self.p, None, 0)
        arr = (Int3 * max(1, n))()
        self.lib.orc_sim_drag_links(self.p, arr, n)
        return [(arr[i].i, arr[i].j, arr[i].k) for i in range(n)]

    def line(self, which, b):
        L = Inf_Ff_store()
        self.lib.orc_sim_line(self.p, which, b, C.byref(L))
        return L

    def markers(self, b):
        m = (C.c_int * 3)()
        self.lib.orc_sim_markers(self.p, b, m)
        return list(m)

    def ring(self, b, slot, k, n):
        out = np.zeros(n)
        self.lib.orc_sim_ring(self.p, b, slot, k, _dptr(out))
        return out

    def set_block_hook(self, fn):
        """fn(area, level, substep) replaces each block's stream-collide (None restores it)."""
        if fn is None:
            self._hook = None
            self.lib.orc_sim_set_block_hook(self.p, None, None)
            return
        self._hook = BLOCK_HOOK(lambda ctx, a, level, i: fn(a, level, i))
        self.lib.orc_sim_set_block_hook(self.p, C.cast(self._hook, C.c_void_p), None)

    def spline_line(self, b):
        self.lib.orc_sim_spline_line(self.p, b)

    def transfer_line(self, b, which):
        self.lib.orc_sim_transfer_line(self.p, b, which)

    def close(self):
        if self.p:
            self.lib.orc_sim_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OraclePeriodic:
    def __init__(self, nx, ny, tau=0.6, rho0=1.0, seed=0x5EED2024, sigma=0.01):
        self.lib = load()
        self.nx, self.ny = nx, ny
        self.p = self.lib.orc_periodic_create(nx, ny, tau, rho0, seed, sigma)

    def run(self, n):
        return self.lib.orc_periodic_run(self.p, n)

    def get(self, field):
        out = np.zeros(self.nx * self.ny)
        assert self.lib.orc_periodic_get(self.p, field_id(field), _dptr(out)) == 0
        return out.reshape(self.ny, self.nx)

    def close(self):
        if self.p:
            self.lib.orc_periodic_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spline_column(g, r):
    lib = load()
    g = np.ascontiguousarray(g, dtype=np.float64)
    Nc = g.size
    M = np.zeros(Nc)
    out = np.zeros(r * (Nc - 1) + 1)
    lib.orc_spline_column(Nc, r, _dptr(g), _dptr(M), _dptr(out))
    return M, out
