"""The C-ABI boundary: library loading, exported symbols, struct layouts of
the Python mirror against the C compiler, and the host-only geometry code
(geo/init_geo.h restated in mbl_geo.c) against the oracle's restatement.
No GPU is touched here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import mbl
from mbl import abi
import oracle

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "mbl.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", src, flags=re.M):
        name = m.group(1)
        if name not in ("defined", "sizeof"):
            names.add(name)
    return names


@pytest.mark.parametrize("fp64", [False, True])
def test_library_exports_every_header_symbol(fp64):
    lib = mbl.load_library(fp64)
    names = header_functions()
    assert len(names) >= 35
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(abi.EXPORTS) <= names


def test_reference_entry_points_present(lib32):
    for n in ("init_General", "init_Fstore", "init_tri_A", "evolution", "sum_N_drag", "fdrag_store",
              "getDrag_block", "error_L2"):
        assert hasattr(lib32, n)


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "mbl.h"
#define S(T) printf(#T " %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  S(D2Q9) S(AREA) S(Inf_Ff_store) S(Ff_Store) S(FM_spline3) S(F_k3) S(Int3) S(P4) S(P3i) S(P4d4i)
  S(mbl_config) S(mbl_geometry) S(mbl_channel_spec) S(mbl_dim3)
  O(D2Q9, F) O(D2Q9, f) O(AREA, base_x) O(AREA, rv_tau) O(Inf_Ff_store, line_F) O(Ff_Store, f_t_1)
  O(FM_spline3, F) O(mbl_config, seed) O(mbl_config, out_dir) O(mbl_geometry, AreaZone) O(mbl_geometry, R)
  return 0; }
"""


@pytest.mark.parametrize("fp64", [False, True])
def test_python_mirror_matches_c_layout(tmp_path, fp64):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    cmd = ["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)]
    if fp64:
        cmd.insert(1, "-DMBL_FP64")
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    py = {"D2Q9": abi.D2Q9, "AREA": abi.AREA, "Inf_Ff_store": abi.Inf_Ff_store, "Ff_Store": abi.Ff_Store,
          "FM_spline3": abi.FM_spline3, "F_k3": abi.F_k3, "Int3": abi.Int3, "P4": abi.P4, "P3i": abi.P3i,
          "P4d4i": abi.P4d4i, "mbl_config": abi.mbl_config, "mbl_geometry": abi.mbl_geometry,
          "mbl_channel_spec": abi.mbl_channel_spec, "mbl_dim3": abi.mbl_dim3}
    for name, T in py.items():
        assert int(got[name]) == C.sizeof(T), name
    for key, val in got.items():
        if "." in key:
            t, f = key.split(".")
            assert getattr(py[t], f).offset == int(val), key


def geo_both(lib, olib, G):
    g = G.g
    out = []
    for fn in (lambda *a: lib.mbl_geo_init(C.byref(g), *a), lambda *a: olib.orc_geo_init(C.byref(g), *a)):
        ad = (abi.P4 * g.number_area)()
        ar = (abi.AREA * g.number_area)()
        lc = (abi.Inf_Ff_store * max(1, g.number_line))()
        lf = (abi.Inf_Ff_store * max(1, g.number_line))()
        lcc = (abi.Inf_Ff_store * max(1, g.number_line2))()
        assert fn(ad, ar, lc, lf, lcc) == 0
        out.append((bytes(ad), bytes(ar), bytes(lc), bytes(lf), bytes(lcc)))
    return out


@pytest.mark.parametrize("name", ["GEO_1_1", "GEO_2_2", "GEO_2_5", "GEO_4_9"])
def test_geo_init_matches_oracle_presets(lib32, olib, name):
    a, b = geo_both(lib32, olib, mbl.Geometry.preset(lib32, name))
    assert a == b


@pytest.mark.parametrize("regions", [[(20, 100, 20, 60)], [(448, 704, 192, 320), (480, 608, 224, 288)],
                                     [(12, 44, 6, 26), (16, 36, 10, 22)], []])
def test_channel_builder_matches_oracle_and_invariants(lib32, olib, regions):
    LX, LY = (2049, 513) if regions and regions[0][1] > 300 else (301, 81)
    G = mbl.Geometry.channel(lib32, LX, LY, regions, R=4.0, CX=40.0, CY=40.0, Re=50.0)
    a, b = geo_both(lib32, olib, G)
    assert a == b
    ad, areas, lc, lf, lcc = G.init(lib32)
    g = G.g
    assert g.LEVEL == len(regions) + 1
    assert g.LTCFFCP[g.LEVEL] == g.LS3CFFCP[g.LEVEL - 1] == g.number_line
    for b_ in range(g.number_line):
        L = lc[b_]
        r = areas[L.areaf_n].m // areas[L.areac_n].m
        assert r == 2 and L.Nf == r * (L.Nc - 1) + 1
    # the whole domain is covered by level 0 blocks and every region by its level
    cover = np.zeros((LY, LX), dtype=int)
    for a_ in range(g.ASP[0], g.ASP[1]):
        p = ad[a_]
        cover[p.ystart:p.yend + 1, p.xstart:p.xend + 1] += 1
    if regions:
        x0, x1, y0, y1 = regions[0]
        cover[y0 + 2:y1 - 1, x0 + 2:x1 - 1] += 1
    assert cover.min() >= 1


def test_geo_init_rejects_tau_one(lib32):
    # nu = 1/6 makes tau_0 = 1: the f_neq rescale (tau_f - 1)/(r (tau_c - 1)) is singular
    U, R = 0.1, 5.0
    G = mbl.Geometry.channel(lib32, 301, 81, [(20, 100, 20, 60)], R=R, CX=60.0, CY=40.0, Re=U * 2 * R * 6.0)
    g = G.g
    ad = (abi.P4 * g.number_area)()
    ar = (abi.AREA * g.number_area)()
    lc = (abi.Inf_Ff_store * g.number_line)()
    lf = (abi.Inf_Ff_store * g.number_line)()
    lcc = (abi.Inf_Ff_store * g.number_line2)()
    assert lib32.mbl_geo_init(C.byref(g), ad, ar, lc, lf, lcc) == -1


def test_error_strings(lib32):
    for code, name in abi.E_NAMES.items():
        s = lib32.mbl_error_string(code).decode()
        assert s


def test_driver_is_plain_c():
    src = open(os.path.join(REPO, "multi-block-lbm-gpu_amd", "driver", "mbl_driver.c")).read()
    assert "hip" not in re.sub(r"mbl_hip", "", src.lower()).replace("hip_", "")
    assert '#include "../../include/mbl.h"' in src


def test_no_oracle_in_product():
    pkg = os.path.join(REPO, "multi-block-lbm-gpu_amd")
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".c", ".h", ".cpp")) or f == "Makefile":
                text = open(os.path.join(root, f), errors="ignore").read()
                assert "oracle" not in text.lower(), os.path.join(root, f)
