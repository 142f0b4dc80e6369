"""The C-ABI library on the CPU: it loads, exports every declared symbol,
and its host-side functions (parameter file, grid guard, geometry) behave as
the reference's driver does.  No GPU compute here."""
import ctypes as C
import io
import os
import re
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import gubingluo_amd as G
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "rtm_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"^\s*#.*$", "", src, flags=re.M)          # preprocessor lines (macros are not symbols)
    # function declarations; a function-pointer member `int (*f)(...)` is not one
    return sorted(set(re.findall(r"\b([A-Za-z_]\w*)\s*\((?!\s*\*)[^;{]*\)\s*;", src)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert "rtm_model" in names and "rtm_real" in names and "rtmh_run" in names
    L = G.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(G.EXPORTED) <= set(names)


def test_version_and_no_device_here():
    assert b"gfx950" in G.lib().rtmh_version()
    assert G.device_count() >= 0


PARFILE = """dt(s) nt tsnap(s)
0.001 2000 0.1
xmax zmax dx dz (m)
3990 1990 10 10
fdom amp alp
20 1 0.05
spx0 spz0 dspx
2000 20 50
offset offsetmax
0 4000
startshot nshot dshot nshotall
1 1 1 1
npml
40
medium_flag ismth
0 0
direct direct0
1.5 0.05
idevice
0
flag_layout
2
mode
0
dataflag
0
sgyflag
0
endianflag
0
rhoflag
0
vp_file
vel.su
"""


def test_parfile_parse():
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "par.txt")
        open(p, "w").write(PARFILE)
        pf = G.Parfile()
        assert G.lib().rtmh_read_parfile(p.encode(), C.byref(pf)) == 0
        assert abs(pf.dt0 - 0.001) < 1e-9 and pf.nt0 == 2000 and pf.npml == 40 and pf.flag_layout == 2
        assert pf.xmax == 3990 and pf.dx0 == 10 and pf.spx0 == 2000 and pf.offsetmax == 4000
        assert pf.vp_file == b"vel.su" and pf.rhoflag == 0 and pf.dataflag == 0
        assert G.lib().rtmh_read_parfile(os.path.join(td, "none").encode(), C.byref(pf)) == -1


def _grid(vp, fdom, dx, nt, dt):
    g = G.Grid()
    nz, nx = vp.shape
    G.lib().rtmh_get_constant(G._fp(np.ascontiguousarray(vp, np.float32)), C.c_float(fdom), C.c_float(dx),
                              C.c_float(dx), C.c_int(nx), C.c_int(nz), C.c_int(nt), C.c_float(dt), C.byref(g))
    return g


def test_get_constant_guards():
    """dx_max = vmin/fdom/7.25 halves the spacing; dt_max = h/(sqrt2 sum|c| vmax)
    shrinks dt (RTM_GPU_velstr.cpp:319-362)."""
    vp = np.full((20, 30), 2000.0, np.float32)
    vp[10:] = 3000.0
    g = _grid(vp, 20.0, 10.0, 2000, 1e-3)
    assert (g.nx, g.nz, g.lx, g.lz, g.nt) == (30, 20, 1, 1, 2000) and abs(g.dt - 1e-3) < 1e-9
    g = _grid(vp, 40.0, 10.0, 2000, 1e-3)                  # 2000/40/7.25 = 6.9 m < 10 m
    assert (g.lx, g.lz, g.nx, g.nz) == (2, 2, 59, 39) and g.dx == 5.0
    dtmax = 5.0 / (np.sqrt(2.0) * 1.4760199 * 3000.0)
    assert abs(g.dt - np.floor(dtmax * 1e4) * 1e-4) < 1e-7
    assert g.nt == int(np.float32(2000 * np.float32(1e-3)) / np.float32(g.dt))


@pytest.mark.parametrize("layout", [1, 2, 3, 4])
def test_geometry_matches_oracle(layout):
    kw = dict(nx=500, nz=100, npml=50, flag_layout=layout, spx0=35.0, dspx=37.5, offset=55.0, offsetmax=1500.0,
              dx=7.5, dz=7.5, nshot=60)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    for ishot in range(1, 61, 3):
        try:
            a = G.geometry(P, ishot)
        except ValueError:
            with pytest.raises(ValueError):
                O.geometry(Po, ishot)
            continue
        b = O.geometry(Po, ishot)
        for f in ("nx1", "nx2", "nxl", "nxp", "nzp", "nsx", "nsz", "nrec", "rcol0", "noffset"):
            assert getattr(a, f) == getattr(b, f), (layout, ishot, f)


def test_selfdoc_banner():
    code = ("import ctypes as C, gubingluo_amd as G\n"
            "pf = G.Parfile(); pf.nshot=200; pf.startshot=1; pf.dshot=1; pf.nt0=4000; pf.dz0=10; pf.dx0=10\n"
            "pf.dt0=0.0008; pf.fdom=15; pf.npml=50; pf.flag_layout=1; pf.vp_file=b'vel.su'\n"
            "g = G.Grid(10.0, 10.0, 0.0008, 1024, 512, 4000, 1, 1, 1500.0, 4500.0)\n"
            "G.lib().rtmh_selfdoc(C.byref(pf), 1024, 512, C.byref(g))\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT).stdout
    assert "Prestack Reverse Time Migration by Bingluo Gu from UPC." in out
    assert " npml      =  50  pml absorb layers number" in out
    assert " Norder    =  20  order of finite difference" in out
    assert " layout:       1  1: Bilateral;  2: full array;  3: Unilateral(r);  4: Unilateral(l)" in out


def test_run_without_gpu_fails_loudly():
    if G.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        G.Context(0)
