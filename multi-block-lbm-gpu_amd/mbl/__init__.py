"""Python host mirror of libmbl_hip's C ABI (include/mbl.h).

`MultiBlockSim` reproduces the reference driver (Multi-block-LBM.cu:19-230):
geometry set-up (geo_init), per-block allocation, init_General, the drag-link
set-up (sum_N_drag / fdrag_store), init_Fstore / init_tri_A and the time loop
evolution -> error_L2 -> getDrag_block. `PeriodicBlock` is the single
periodic block of the bandwidth benchmark. Everything runs through the C ABI
of the HIP library; the library must be built (make -C multi-block-lbm-gpu_amd)
and there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import (AREA, D2Q9, F_k3, FM_spline3, Ff_Store, Inf_Ff_store, Int3, P3i, P4, mbl_case_spec,
                  mbl_channel_spec, mbl_config, mbl_dim3, mbl_geometry)

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")

_libs = {}


class MblError(RuntimeError):
    pass


def library_path(fp64=False):
    return os.path.join(LIB_DIR, "libmbl_hip64.so" if fp64 else "libmbl_hip.so")


def load_library(fp64=False):
    """Loads libmbl_hip(.so|64.so); raises if it has not been built."""
    key = bool(fp64)
    if key in _libs:
        return _libs[key]
    path = library_path(fp64)
    if not os.path.exists(path):
        raise MblError(f"{path} is not built: run `make -C {PKG_DIR}` (or __graft_entry__.build())")
    lib = C.CDLL(path)
    vp, ip, dp, fp = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_float)
    sig = {
        "mbl_config_default": (None, [C.POINTER(mbl_config)]),
        "mbl_handle_create": (ip, [C.POINTER(vp), C.POINTER(mbl_config)]),
        "mbl_handle_destroy": (None, [vp]),
        "mbl_handle_make_current": (None, [vp]),
        "mbl_handle_current": (vp, []),
        "mbl_get_last_error": (ip, [vp]),
        "mbl_error_string": (C.c_char_p, [ip]),
        "mbl_synchronize": (ip, [vp]),
        "mbl_flush_diagnostics": (ip, [vp]),
        "mbl_set_stream": (ip, [vp, vp]),
        "mbl_malloc": (vp, [C.c_size_t]),
        "mbl_free": (None, [vp]),
        "mbl_memcpy_h2d": (ip, [vp, vp, C.c_size_t]),
        "mbl_memcpy_d2h": (ip, [vp, vp, C.c_size_t]),
        "mbl_memcpy_d2d": (ip, [vp, vp, C.c_size_t]),
        "mbl_memset_d": (ip, [vp, ip, C.c_size_t]),
        "mbl_plane_bytes": (C.c_size_t, [vp, ip, ip]),
        "mbl_block_alloc": (ip, [vp, C.POINTER(AREA), C.POINTER(D2Q9), C.POINTER(vp), C.POINTER(vp)]),
        "mbl_block_free": (None, [C.POINTER(D2Q9), vp, vp]),
        "mbl_read_distribution": (ip, [vp, C.POINTER(D2Q9), ip, ip, C.c_size_t, dp]),
        "mbl_write_distribution": (ip, [vp, C.POINTER(D2Q9), ip, ip, C.c_size_t, dp]),
        "mbl_geo_init": (ip, [C.POINTER(mbl_geometry), C.POINTER(P4), C.POINTER(AREA), C.POINTER(Inf_Ff_store),
                              C.POINTER(Inf_Ff_store), C.POINTER(Inf_Ff_store)]),
        "mbl_geometry_preset": (ip, [C.c_char_p, C.POINTER(mbl_geometry)]),
        "mbl_geometry_channel": (ip, [C.POINTER(mbl_channel_spec), C.POINTER(mbl_geometry), C.POINTER(P4),
                                      C.POINTER(P3i), C.POINTER(P3i)]),
        "init_General": (None, [mbl_dim3, mbl_dim3, AREA, D2Q9, vp, vp]),
        "init_Fstore": (None, [mbl_dim3, mbl_dim3, AREA, D2Q9, Inf_Ff_store, Ff_Store]),
        "init_tri_A": (None, [mbl_dim3, mbl_dim3, AREA, D2Q9, Inf_Ff_store, FM_spline3]),
        "evolution": (None, [ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), ip, mbl_dim3,
                             C.POINTER(mbl_dim3), C.POINTER(AREA), C.POINTER(D2Q9), C.POINTER(D2Q9),
                             C.POINTER(Inf_Ff_store), C.POINTER(Inf_Ff_store), C.POINTER(Inf_Ff_store),
                             C.POINTER(Ff_Store), C.POINTER(FM_spline3), vp]),
        "sum_N_drag": (ip, [mbl_dim3, mbl_dim3, AREA, D2Q9]),
        "fdrag_store": (None, [AREA, D2Q9, vp, D2Q9, ip]),
        "getDrag_block": (None, [ip, vp, AREA, D2Q9, dp, ip]),
        "mbl_evolution_step_fused": (ip, [vp]),
        "mbl_block_moments": (ip, [vp, C.POINTER(AREA), C.POINTER(D2Q9), dp]),
        "mbl_evolution_step": (ip, [vp, ip, ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), mbl_dim3,
                                    C.POINTER(mbl_dim3), C.POINTER(AREA), C.POINTER(D2Q9), C.POINTER(D2Q9),
                                    C.POINTER(Inf_Ff_store), C.POINTER(Inf_Ff_store), C.POINTER(Inf_Ff_store),
                                    C.POINTER(Ff_Store), C.POINTER(FM_spline3), C.POINTER(vp), C.POINTER(vp), vp, ip,
                                    dp]),
        "error_L2": (None, [ip, ip, mbl_dim3, C.POINTER(mbl_dim3), C.POINTER(AREA), C.POINTER(D2Q9),
                            C.POINTER(vp), C.POINTER(vp), vp, vp, vp, vp]),
        "mbl_last_diagnostics": (ip, [vp, dp]),
        "mbl_diagnostics_series": (ip, [vp, ip, ip, C.POINTER(ip), dp, dp]),
        "mbl_diagnostics_clear": (None, [vp]),
        "mbl_divergence_step": (ip, [vp]),
        "mbl_cycle_launches": (ip, [vp, ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)]),
        "mbl_cycle_positions": (ip, [vp, ip]),
        "mbl_set_cycle_fork": (ip, [vp, ip]),
        "mbl_error_L2_noalloc": (ip, [vp, ip, C.POINTER(AREA), C.POINTER(D2Q9), C.POINTER(vp), C.POINTER(vp), vp]),
        "mbl_step_block": (ip, [vp, C.POINTER(AREA), C.POINTER(D2Q9), ip]),
        "mbl_set_step_pairs": (ip, [vp, ip]),
        "mbl_step_block_launches": (ip, [vp, C.POINTER(AREA), ip]),
        "mbl_timer_begin": (ip, [vp]),
        "mbl_timer_end": (ip, [vp, dp]),
        "mbl_set_tuning": (ip, [vp, ip, ip]),
        "mbl_poiseuille_l2": (ip, [vp, ip, C.POINTER(AREA), C.POINTER(D2Q9), dp]),
        "mbl_set_persistent": (ip, [vp, ip]),
        "mbl_batch_create": (ip, [vp, ip, C.POINTER(mbl_case_spec), C.POINTER(vp)]),
        "mbl_batch_destroy": (None, [vp]),
        "mbl_batch_step": (ip, [vp, ip]),
        "mbl_batch_set_persistent": (ip, [vp, ip]),
        "mbl_batch_node_updates": (C.c_double, [vp]),
        "mbl_batch_forces": (ip, [vp, ip, dp, dp]),
        "mbl_batch_residual": (ip, [vp, dp]),
        "mbl_batch_autotune": (ip, [vp, ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), dp]),
        "mbl_batch_block": (ip, [vp, ip, ip, C.POINTER(AREA), C.POINTER(D2Q9)]),
        "mbl_batch_raster": (ip, [vp, ip, ip, fp]),
        "mbl_batch_init_from_raster": (ip, [vp, ip, ip, fp]),
        "mbl_bandwidth_probe": (ip, [vp, C.c_size_t, ip, dp]),
        "mbl_bandwidth_probe_planes": (ip, [vp, C.c_size_t, ip, dp]),
        "output": (None, [ip, C.POINTER(AREA), C.POINTER(D2Q9), C.POINTER(D2Q9)]),
        "mbl_output": (ip, [vp, ip, ip, C.POINTER(AREA), C.POINTER(D2Q9), C.POINTER(D2Q9)]),
        "mbl_checkpoint_save": (ip, [vp, C.c_char_p, C.c_longlong, ip, C.POINTER(AREA), C.POINTER(D2Q9),
                                     C.POINTER(vp), C.POINTER(vp), ip, C.POINTER(Ff_Store)]),
        "mbl_checkpoint_load": (ip, [vp, C.c_char_p, C.POINTER(C.c_longlong), ip, C.POINTER(AREA), C.POINTER(D2Q9),
                                     C.POINTER(vp), C.POINTER(vp), ip, C.POINTER(Ff_Store)]),
        "mbl_net_create": (ip, [vp, ip, ip, ip, ip, ip, ip, C.c_ulonglong, C.POINTER(vp)]),
        "mbl_net_destroy": (None, [vp]),
        "mbl_net_num_params": (C.c_longlong, [vp]),
        "mbl_net_params": (ip, [vp, fp, ip]),
        "mbl_net_forward": (ip, [vp, fp, ip, fp]),
        "mbl_net_loss_grad": (ip, [vp, fp, fp, ip, dp, fp]),
        "mbl_net_train": (ip, [vp, fp, fp, ip, ip, ip, C.c_float, C.c_ulonglong, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib.fp64 = key
    lib.real = np.float64 if key else np.float32
    _libs[key] = lib
    return lib


def check(lib, rc, what=""):
    if rc:
        raise MblError(f"{what}: {abi.E_NAMES.get(rc, rc)} ({rc})")


def default_config(lib, **kw):
    cfg = mbl_config()
    lib.mbl_config_default(C.byref(cfg))
    for k, v in kw.items():
        if k == "mi":
            for i, x in enumerate(v):
                cfg.mi[i] = x
        elif k == "U0":
            cfg.U0[0], cfg.U0[1] = v
        elif k == "out_dir":
            cfg.out_dir = v.encode() if isinstance(v, str) else v
        else:
            setattr(cfg, k, v)
    return cfg


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------
class Geometry:
    """A runtime layout (mbl_geometry) plus the cylinder and channel constants."""

    def __init__(self, g, keep=(), CX=None, CY=None, bc_mode=abi.BC_REFERENCE, init_mode=abi.INIT_UNIFORM):
        self.g = g
        self._keep = keep
        self.CX, self.CY = CX, CY
        self.bc_mode = bc_mode
        self.init_mode = init_mode

    @classmethod
    def preset(cls, lib, name):
        g = mbl_geometry()
        check(lib, lib.mbl_geometry_preset(name.encode(), C.byref(g)), f"preset {name}")
        return cls(g, CX=160.0, CY=160.0)

    @classmethod
    def channel(cls, lib, LX, LY, regions, R, CX, CY, Re, U=0.1):
        """Nested-refinement channel; regions[l-1] = (x0, x1, y0, y1) of level l."""
        s = mbl_channel_spec()
        s.LX, s.LY, s.levels = LX, LY, len(regions) + 1
        for l, r in enumerate(regions, start=1):
            s.region[l] = P4(*r)
        s.Re, s.U, s.R = Re, U, R
        nb = 4 * s.levels
        ad = (P4 * nb)()
        cf = (P3i * (8 * s.levels))()
        cc = (P3i * (8 * s.levels))()
        g = mbl_geometry()
        check(lib, lib.mbl_geometry_channel(C.byref(s), C.byref(g), ad, cf, cc), "channel geometry")
        return cls(g, keep=(s, ad, cf, cc), CX=CX, CY=CY, bc_mode=abi.BC_CHANNEL, init_mode=abi.INIT_REST)

    def init(self, lib):
        """geo_init: (AreaDefine, Areas, Line_Inf_C, Line_Inf_F, Line_Inf_CC)."""
        g = self.g
        ad = (P4 * g.number_area)()
        areas = (AREA * g.number_area)()
        lc = (Inf_Ff_store * max(1, g.number_line))()
        lf = (Inf_Ff_store * max(1, g.number_line))()
        lcc = (Inf_Ff_store * max(1, g.number_line2))()
        check(lib, lib.mbl_geo_init(C.byref(g), ad, areas, lc, lf, lcc), "geo_init")
        return ad, areas, lc, lf, lcc


def c1_geometry(lib):
    """Two-level 300x80-cell channel, fine block 80x40 cells, cylinder D=10 at (60, 40), Re=20."""
    return Geometry.channel(lib, 301, 81, [(20, 100, 20, 60)], R=5.0, CX=60.0, CY=40.0, Re=20.0)


def c2b_geometry(lib):
    """Three-level 2048x512-cell channel, cylinder D=16 coarse (64 finest) at (512, 256.25), Re=100.

    The quarter-coarse-cell (one finest cell) offset from the centre line breaks the mirror
    symmetry the same way at every precision, so vortex shedding does not depend on round-off
    (16-bit planes round the round-off asymmetry away). The flow starts from the developed
    inlet profile, which avoids the acoustic ringing of an impulsive start in a channel whose
    ends reflect sound."""
    g = Geometry.channel(lib, 2049, 513, [(448, 704, 192, 320), (480, 608, 224, 288)], R=8.0,
                         CX=512.0, CY=256.25, Re=100.0)
    g.init_mode = abi.INIT_PROFILE
    return g


# ---------------------------------------------------------------------------
# multi-block run (the reference driver, Multi-block-LBM.cu:19-230)
# ---------------------------------------------------------------------------
class MultiBlockSim:
    def __init__(self, lib, geom, storage=abi.STORE_NATIVE, use_graph=1, sync_diag=0, write_files=0,
                 out_dir="", flush_interval=1000, rho0=1.0, U0=(0.1, 0.0), init_mode=None, bc_mode=None,
                 tuning=None, fused=True):
        self.lib = lib
        self.fused = fused
        self.geom = geom
        g = geom.g
        self.real = lib.real
        self.ad, self.areas, self.lc, self.lf, self.lcc = geom.init(lib)
        self.na, self.nl, self.nl2 = g.number_area, g.number_line, g.number_line2
        cfg = default_config(lib, LEVEL=g.LEVEL, number_line2=g.number_line2, rho0=rho0, U0=U0, LX=g.LX, LY=g.LY,
                             R=g.R, CX=geom.CX, CY=geom.CY,
                             bc_mode=geom.bc_mode if bc_mode is None else bc_mode,
                             init_mode=geom.init_mode if init_mode is None else init_mode,
                             storage=storage, use_graph=use_graph, sync_diag=sync_diag, write_files=write_files,
                             out_dir=out_dir, flush_interval=flush_interval, number_area=g.number_area)
        for i in range(g.LEVEL):
            cfg.mi[i] = g.mi[i]
        self.cfg = cfg
        h = C.c_void_p()
        check(lib, lib.mbl_handle_create(C.byref(h), C.byref(cfg)), "handle")
        self.h = h
        if tuning:
            check(lib, lib.mbl_set_tuning(h, *tuning), "tuning")
        self.pps_d = (D2Q9 * self.na)()
        self.pps_h = (D2Q9 * self.na)()
        self.upx = (C.c_void_p * self.na)()
        self.upy = (C.c_void_p * self.na)()
        for a in range(self.na):
            ux, uy = C.c_void_p(), C.c_void_p()
            check(lib, lib.mbl_block_alloc(h, C.byref(self.areas[a]), C.byref(self.pps_d[a]), C.byref(ux),
                                           C.byref(uy)), "block alloc")
            self.upx[a], self.upy[a] = ux.value, uy.value
        rb = np.dtype(self.real).itemsize
        self._dev = []
        self.t3 = (Ff_Store * max(1, self.nl))()
        self.s3 = (FM_spline3 * max(1, self.nl))()
        for b in range(self.nl):
            t = self.t3[b]
            t.marker[0], t.marker[1], t.marker[2] = 2, 1, 0
            t.N = self.lc[b].Nf
            for slot in (t.f_t1, t.f_t, t.f_t_1):
                for k in range(abi.Q):
                    slot.F[k] = self._alloc(t.N * rb)
            s = self.s3[b]
            s.N = self.lc[b].Nc
            s.l, s.m, s.u = self._alloc(s.N * rb), self._alloc(s.N * rb), self._alloc(s.N * rb)
            s.Varibs_M = self._alloc(12 * s.N * rb)
            s.rho, s.ux, s.uy = self._alloc(s.N * rb), self._alloc(s.N * rb), self._alloc(s.N * rb)
            for k in range(abi.Q):
                s.F[k] = self._alloc(s.N * rb)
        self.block = mbl_dim3(16, 16, 1)
        self.grid = (mbl_dim3 * self.na)()
        for a in range(self.na):
            self.grid[a] = mbl_dim3((self.areas[a].nx + 15) // 16, (self.areas[a].ny + 15) // 16, 1)
        lib.mbl_handle_make_current(h)
        for a in range(self.na):
            lib.init_General(self.grid[a], self.block, self.areas[a], self.pps_d[a], self.upx[a], self.upy[a])
        last = self.na - 1
        self.n_drag = lib.sum_N_drag(self.block, self.grid[last], self.areas[last], self.pps_d[last])
        self.d_links = self._alloc(max(1, self.n_drag) * C.sizeof(Int3))
        lib.fdrag_store(self.areas[last], self.pps_h[last], self.d_links, self.pps_d[last], self.n_drag)
        for b in range(self.nl):
            lc = self.lc[b]
            lib.init_Fstore(self.block, self.block, self.areas[lc.areaf_n], self.pps_d[lc.areaf_n], lc, self.t3[b])
            lib.init_tri_A(self.block, self.block, self.areas[lc.areac_n], self.pps_d[lc.areac_n], lc, self.s3[b])
        self.asp = (C.c_int * (g.LEVEL + 1))(*[g.ASP[i] for i in range(g.LEVEL + 1)])
        self.ls3 = (C.c_int * (g.LEVEL + 1))(*[g.LS3CFFCP[i] for i in range(g.LEVEL + 1)])
        self.ltc = (C.c_int * (g.LEVEL + 1))(*[g.LTCFFCP[i] for i in range(g.LEVEL + 1)])
        self.lccp = (C.c_int * (g.LEVEL + 1))(*[g.LCCP[i] for i in range(g.LEVEL + 1)])
        self.t = 0
        self.sync()

    def _alloc(self, nbytes):
        p = self.lib.mbl_malloc(max(8, nbytes))
        if not p:
            raise MblError("mbl_malloc failed")
        self._dev.append(p)
        return p

    def sync(self):
        rc = self.lib.mbl_synchronize(self.h)
        check(self.lib, rc, "synchronize")

    def last_error(self):
        return self.lib.mbl_get_last_error(self.h)

    def evolve(self):
        self.lib.mbl_handle_make_current(self.h)
        self.lib.evolution(self.t, self.asp, self.ls3, self.ltc, self.lccp, 0, self.block, self.grid, self.areas,
                           self.pps_h, self.pps_d, self.lc, self.lf, self.lcc, self.t3, self.s3, self.h)

    def step(self, n=1, diagnostics=True, fused=None):
        """n coarse steps of the reference loop: evolution, then (diagnostics) error_L2 and
        getDrag_block. fused (default: the sim's): the n steps in one mbl_evolution_step call."""
        lib = self.lib
        lib.mbl_handle_make_current(self.h)
        last = self.na - 1
        if diagnostics and (self.fused if fused is None else fused):
            rc = lib.mbl_evolution_step(self.h, self.t, n, self.asp, self.ls3, self.ltc, self.lccp, self.block,
                                        self.grid, self.areas, self.pps_h, self.pps_d, self.lc, self.lf, self.lcc,
                                        self.t3, self.s3, self.upx, self.upy, self.d_links, self.n_drag, None)
            self.t += n
            if rc:
                raise MblError(f"step failed: {abi.E_NAMES.get(rc, rc)}")
            return
        for _ in range(n):
            self.evolve()
            if diagnostics:
                lib.error_L2(self.t, self.na, self.block, self.grid, self.areas, self.pps_d, self.upx, self.upy,
                             None, None, None, None)
                if self.n_drag > 0:
                    lib.getDrag_block(self.t, self.d_links, self.areas[last], self.pps_d[last], None, self.n_drag)
            self.t += 1
        rc = self.last_error()
        if rc:
            raise MblError(f"step failed: {abi.E_NAMES.get(rc, rc)}")

    def flush(self):
        check(self.lib, self.lib.mbl_flush_diagnostics(self.h), "flush")

    def output(self, n=None):
        """steady_<n>.dat (Tecplot point file of x, y, solid, rho, U, V per block) in out_dir."""
        check(self.lib, self.lib.mbl_output(self.h, self.t if n is None else n, self.na, self.areas, None,
                                            self.pps_d), "output")

    def checkpoint(self, path):
        """Saves the whole time-stepping state (planes, macros, previous velocities, rings, markers, step)."""
        check(self.lib, self.lib.mbl_checkpoint_save(self.h, str(path).encode(), self.t, self.na, self.areas,
                                                     self.pps_d, self.upx, self.upy, self.nl, self.t3),
              "checkpoint save")

    def restore(self, path):
        """Loads a checkpoint of the same layout and storage written by checkpoint()."""
        step = C.c_longlong()
        check(self.lib, self.lib.mbl_checkpoint_load(self.h, str(path).encode(), C.byref(step), self.na, self.areas,
                                                     self.pps_d, self.upx, self.upy, self.nl, self.t3),
              "checkpoint load")
        self.t = step.value

    def poiseuille_l2(self):
        """Velocity L2 distance from the analytic channel profile (FLUID nodes, 1/m weights)."""
        out = C.c_double()
        check(self.lib, self.lib.mbl_poiseuille_l2(self.h, self.na, self.areas, self.pps_d, C.byref(out)),
              "poiseuille_l2")
        return out.value

    def series(self):
        """{'t','residual'} and {'t','Cd','Cl'} recorded so far (after a flush)."""
        self.flush()
        lib, h = self.lib, self.h
        n0 = lib.mbl_diagnostics_series(h, 0, 0, None, None, None)
        if n0 < 0:
            check(lib, n0, "diagnostics series")
        t0 = (C.c_int * max(1, n0))()
        r0 = np.zeros(max(1, n0))
        lib.mbl_diagnostics_series(h, 0, n0, t0, r0.ctypes.data_as(C.POINTER(C.c_double)), None)
        n1 = lib.mbl_diagnostics_series(h, 1, 0, None, None, None)
        if n1 < 0:
            check(lib, n1, "diagnostics series")
        t1 = (C.c_int * max(1, n1))()
        cd, cl = np.zeros(max(1, n1)), np.zeros(max(1, n1))
        lib.mbl_diagnostics_series(h, 1, n1, t1, cd.ctypes.data_as(C.POINTER(C.c_double)),
                                   cl.ctypes.data_as(C.POINTER(C.c_double)))
        return ({"t": np.array(t0[:n0]), "residual": r0[:n0]},
                {"t": np.array(t1[:n1]), "Cd": cd[:n1], "Cl": cl[:n1]})

    def shape(self, a):
        return self.areas[a].ny, self.areas[a].nx

    def read(self, a, field):
        """field: 'solid', 'x', 'y', 'rho', 'ux', 'uy', 'upx', 'upy', ('F', k) or ('f', k)."""
        self.sync()
        ny, nx = self.shape(a)
        n = nx * ny
        pp = self.pps_d[a]
        if isinstance(field, tuple):
            which, k = field
            out = np.zeros(n)
            rc = self.lib.mbl_read_distribution(self.h, C.byref(pp), 1 if which == "f" else 0, k, n,
                                                out.ctypes.data_as(C.POINTER(C.c_double)))
            check(self.lib, rc, "read distribution")
            return out.reshape(ny, nx)
        if field == "solid":
            out = np.zeros(n, dtype=np.int32)
            src = pp.solid
        else:
            out = np.zeros(n, dtype=self.real)
            src = {"x": pp.x, "y": pp.y, "rho": pp.rho, "ux": pp.ux, "uy": pp.uy,
                   "upx": self.upx[a], "upy": self.upy[a]}[field]
        check(self.lib, self.lib.mbl_memcpy_d2h(out.ctypes.data, src, out.nbytes), "d2h")
        return out.reshape(ny, nx)

    def write_F(self, a, k, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        rc = self.lib.mbl_write_distribution(self.h, C.byref(self.pps_d[a]), 0, k, v.size,
                                             v.ctypes.data_as(C.POINTER(C.c_double)))
        check(self.lib, rc, "write distribution")

    def close(self):
        if getattr(self, "h", None):
            self.lib.mbl_handle_destroy(self.h)
            for a in range(self.na):
                self.lib.mbl_block_free(C.byref(self.pps_d[a]), self.upx[a], self.upy[a])
            for p in self._dev:
                self.lib.mbl_free(p)
            self._dev = []
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# single periodic block (bandwidth benchmark workload)
# ---------------------------------------------------------------------------
class PeriodicBlock:
    """One periodic nx x ny block: rho=1, (u, v) ~ N(0, sigma^2) from `seed`, tau."""

    def __init__(self, lib, nx, ny, tau=0.6, storage=abi.STORE_NATIVE, seed=0x5EED2024, sigma=0.01, tuning=None):
        self.lib = lib
        cfg = default_config(lib, LEVEL=1, number_line2=0, bc_mode=abi.BC_PERIODIC, init_mode=abi.INIT_RANDOM,
                             storage=storage, seed=seed, sigma=sigma, LX=nx, LY=ny, R=0.0, CX=-1e30, CY=-1e30,
                             U0=(0.0, 0.0), write_files=0)
        cfg.mi[0] = 1
        self.cfg = cfg
        h = C.c_void_p()
        check(lib, lib.mbl_handle_create(C.byref(h), C.byref(cfg)), "handle")
        self.h = h
        if tuning:
            check(lib, lib.mbl_set_tuning(h, *tuning), "tuning")
        self.area = AREA(1, nx, ny, 0.0, 0.0, tau, 1.0 / tau)
        self.pp = D2Q9()
        check(lib, lib.mbl_block_alloc(h, C.byref(self.area), C.byref(self.pp), None, None), "block alloc")
        lib.mbl_handle_make_current(h)
        d = mbl_dim3(16, 16, 1)
        lib.init_General(d, d, self.area, self.pp, None, None)
        check(lib, lib.mbl_synchronize(h), "init")

    def step(self, n):
        check(self.lib, self.lib.mbl_step_block(self.h, C.byref(self.area), C.byref(self.pp), n), "step")

    def launches(self, n):
        """Kernel launches of step(n) on this block (two steps per launch where it can)."""
        return self.lib.mbl_step_block_launches(self.h, C.byref(self.area), n)

    def timed_steps(self, n):
        """Runs n steps between HIP events on the handle's stream; returns milliseconds."""
        lib = self.lib
        check(lib, lib.mbl_timer_begin(self.h), "timer")
        self.step(n)
        ms = C.c_double()
        check(lib, lib.mbl_timer_end(self.h, C.byref(ms)), "timer")
        return ms.value

    def sync(self):
        check(self.lib, self.lib.mbl_synchronize(self.h), "synchronize")

    def moments(self):
        """Sum of f, f e_x, f e_y over every node and direction (mass, momentum), and the count
        of non-finite entries, from a device reduction (FP64)."""
        out = np.zeros(4)
        check(self.lib, self.lib.mbl_block_moments(self.h, C.byref(self.area), C.byref(self.pp),
                                                   out.ctypes.data_as(C.POINTER(C.c_double))), "moments")
        return out

    def read_F(self, k, which="F"):
        n = self.area.nx * self.area.ny
        out = np.zeros(n)
        rc = self.lib.mbl_read_distribution(self.h, C.byref(self.pp), 1 if which == "f" else 0, k, n,
                                            out.ctypes.data_as(C.POINTER(C.c_double)))
        check(self.lib, rc, "read")
        return out.reshape(self.area.ny, self.area.nx)

    def close(self):
        if getattr(self, "h", None):
            self.lib.mbl_block_free(C.byref(self.pp), None, None)
            self.lib.mbl_handle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# batched independent cases
# ---------------------------------------------------------------------------
def case_spec(Dc, blockage, Re, U=0.1):
    """Two-level cylinder case in coarse units: diameter Dc (an integer number of
    coarse cells, 2*Dc finest), channel height Dc/blockage, length 20 Dc, refined
    block 4 Dc x 2 Dc with the cylinder one diameter behind its upstream edge."""
    H = int(round(Dc / blockage))
    LX, LY = 20 * Dc + 1, H + 1
    yc = (LY - 1) // 2
    region = P4(3 * Dc, 7 * Dc, yc - Dc, yc + Dc)
    return mbl_case_spec(LX, LY, region, 0.5 * Dc, 4.0 * Dc, float(yc), float(Re), float(U))


def c4_specs(n=256, seed=2024):
    """Seeded sweep: Re ~ U[40, 200], D in {10, 16, 24, 32, 40} finest cells,
    blockage D/H ~ U[0.1, 0.3]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Dfine = int(rng.choice([10, 16, 24, 32, 40]))
        out.append(case_spec(Dfine // 2, float(rng.uniform(0.1, 0.3)), float(rng.uniform(40.0, 200.0))))
    return out


class Batch:
    """Many independent two-level cases advanced by shared launches (mbl_batch_*), or with
    persistent=True by one launch per step() call, one workgroup per case."""

    def __init__(self, lib, specs, storage=abi.STORE_NATIVE, use_graph=1, persistent=False):
        self.lib = lib
        cfg = default_config(lib, LEVEL=2, number_line2=8, bc_mode=abi.BC_CHANNEL, init_mode=abi.INIT_REST,
                             storage=storage, use_graph=use_graph, write_files=0)
        cfg.mi[0], cfg.mi[1] = 1, 2
        self.cfg = cfg
        h = C.c_void_p()
        check(lib, lib.mbl_handle_create(C.byref(h), C.byref(cfg)), "handle")
        self.h = h
        self.specs = (mbl_case_spec * len(specs))(*specs)
        self.n = len(specs)
        b = C.c_void_p()
        check(lib, lib.mbl_batch_create(h, self.n, self.specs, C.byref(b)), "batch create")
        self.b = b
        if persistent:
            self.set_persistent(True)

    def set_persistent(self, on):
        check(self.lib, self.lib.mbl_batch_set_persistent(self.b, 1 if on else 0), "batch persistent")

    def step(self, n):
        check(self.lib, self.lib.mbl_batch_step(self.b, n), "batch step")

    def sync(self):
        check(self.lib, self.lib.mbl_synchronize(self.h), "synchronize")

    def residual(self):
        """error_L2 of every case since the previous call (or since creation / the last restart)."""
        out = np.zeros(self.n)
        check(self.lib, self.lib.mbl_batch_residual(self.b, out.ctypes.data_as(C.POINTER(C.c_double))),
              "batch residual")
        return out

    def node_updates(self):
        return self.lib.mbl_batch_node_updates(self.b)

    def forces(self, nt):
        cd = np.zeros(self.n * nt)
        cl = np.zeros(self.n * nt)
        m = self.lib.mbl_batch_forces(self.b, nt, cd.ctypes.data_as(C.POINTER(C.c_double)),
                                      cl.ctypes.data_as(C.POINTER(C.c_double)))
        if m < 0:
            check(self.lib, m, "batch forces")
        return cd.reshape(self.n, nt)[:, :m], cl.reshape(self.n, nt)[:, :m]

    def block(self, c, q):
        """(AREA, D2Q9) of block q (0..3 coarse, 4 refined) of case c."""
        a, pp = AREA(), D2Q9()
        check(self.lib, self.lib.mbl_batch_block(self.b, c, q, C.byref(a), C.byref(pp)), "batch block")
        return a, pp

    def read(self, c, q, field):
        """Block q of case c: 'F0'..'F8' (decoded to FP64) or a macro field ('rho', 'ux', 'uy', 'solid') as [ny, nx]."""
        a, pp = self.block(c, q)
        n = a.nx * a.ny
        self.sync()
        if field.startswith("F"):
            out = np.zeros(n)
            check(self.lib, self.lib.mbl_read_distribution(self.h, C.byref(pp), 0, int(field[1:]), n,
                                                           out.ctypes.data_as(C.POINTER(C.c_double))), "read F")
        else:
            dt = np.int32 if field == "solid" else self.lib.real
            out = np.zeros(n, dt)
            check(self.lib, self.lib.mbl_memcpy_d2h(out.ctypes.data, getattr(pp, field), out.nbytes), "read field")
        return out.reshape(a.ny, a.nx)

    def raster(self, H, W):
        """(rho, ux, uy) of every case on an H x W raster of its channel: [n, H, W, 3] float32."""
        out = np.zeros((self.n, H, W, 3), np.float32)
        check(self.lib, self.lib.mbl_batch_raster(self.b, H, W, out.ctypes.data_as(C.POINTER(C.c_float))), "raster")
        return out

    def init_from_raster(self, fields):
        """Restarts every case from [n, H, W, 3] (rho, ux, uy) rasters."""
        f = np.ascontiguousarray(fields, np.float32)
        if f.ndim != 4 or f.shape[0] != self.n or f.shape[3] != 3:
            raise ValueError(f"fields must be [{self.n}, H, W, 3], got {f.shape}")
        check(self.lib, self.lib.mbl_batch_init_from_raster(self.b, f.shape[1], f.shape[2],
                                                            f.ctypes.data_as(C.POINTER(C.c_float))), "init raster")

    def autotune(self, trial_steps=10):
        v, t, g, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(self.lib, self.lib.mbl_batch_autotune(self.b, trial_steps, C.byref(v), C.byref(t), C.byref(g),
                                                    C.byref(ms)), "autotune")
        return {"vx": v.value, "tpb": t.value, "groups": g.value, "ms_per_step": ms.value}

    def close(self):
        if getattr(self, "b", None):
            self.lib.mbl_batch_destroy(self.b)
            self.b = None
        if getattr(self, "h", None):
            self.lib.mbl_handle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
