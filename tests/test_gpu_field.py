"""Field-data migration on the GPU: recorded gathers (source at sx,
receivers from the offset to the far end on its right, windows not on the
layout grid) migrated by the engine against the kernel-order oracle shot by
shot, and the `rtm_gpu` driver with dataflag = 1 (gather scan, trace read,
f-k and sinc resampling, rtm_real) against the same chain through the
library's entry points."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import gubingluo_amd as G
from oracle import field as F
from oracle import oracle as O
from test_field_data import gather_bytes
from test_gpu_driver import EXE, read_su, write_su

pytestmark = pytest.mark.gpu


def model(nz, nx, seed):
    rng = np.random.default_rng(seed)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[nz // 2:] = 2600.0
    vp += rng.uniform(-50, 50, (nz, nx)).astype(np.float32)
    rho = (2000.0 + rng.uniform(-100, 100, (nz, nx))).astype(np.float32)
    return vp, rho


@pytest.mark.parametrize("medium_flag,ismth", [(0, 0), (1, 2)])
def test_gather_migration_vs_oracle(gpu_ctx, medium_flag, ismth):
    """with smoothing a gather is muted with the smoothed velocities
    (rtm_real.cu:370-403), unlike rtm_model's data"""
    nz, nx, nt = 50, 200, 260
    vp, rho = model(nz, nx, 21)
    fields = [(123.0, 0.0, 600.0), (487.5, 35.0, 905.0), (1000.0, 20.0, 700.0)]
    kw = dict(nshot=3, startshot=1, dshot=1, mode=1, flag_layout=3, nx=nx, nz=nz, nt=nt, dx=10.0, dz=10.0,
              dt=1e-3, npml=32, fdom=20.0, spz0=20.0, direct=1.5, direct0=0.05, medium_flag=medium_flag,
              ismth=ismth)
    P = G.make_params(**kw)
    rng = np.random.default_rng(22)
    geoms = [G.gather_geometry(P, k + 1, f) for k, f in enumerate(fields)]
    recs = [rng.standard_normal((nt, g.nrec)).astype(np.float32) for g in geoms]
    print(f"[field] 3 gathers 200x50 nt 260 medium_flag {medium_flag} ismth {ismth}", flush=True)
    calls = []
    cb = G.SEIS_WAIT(lambda user, k: calls.append(k) or 0)
    r = gpu_ctx.run(P, vp, rho, want_shots=True, seis_in=recs, shot_field=fields, seis_wait=cb)
    assert calls == [0, 1, 2]                  # each record waited for once, in job order
    for k, f in enumerate(fields):
        Po = O.make_params(**dict(kw, field=1, fsx=f[0], foffset=f[1], foffsetmax=f[2]))
        o = O.run_shot(Po, k + 1, vp, rho, seis_in=recs[k], korder=True)
        assert o.geom.nxl == geoms[k].nxl and o.geom.rcol0 == geoms[k].rcol0
        assert np.array_equal(r.shot_g1[k], o.g1), k
        assert np.array_equal(r.shot_illum[k], o.illum), k
    assert np.abs(r.image1).max() > 0


PARFILE = """dt(s) nt tsnap(s)
0.002 {nt0} 0.1
xmax zmax dx dz (m)
{xmax} {zmax} 20 10
fdom amp alp
20 1 0.05
spx0 spz0 dspx
0 20 0
offset offsetmax
0 0
startshot nshot dshot nshotall
1 3 1 3
npml
32
medium_flag ismth
0 0
direct direct0
1.5 0.05
idevice
0
flag_layout
3
mode
1
dataflag
1
sgyflag
1
endianflag
0
rhoflag
1
vp_file
vel.su
shot_gather
gather.sgy
rho_file
rho.su
"""


def test_rtm_gpu_driver_field_data(gpu_ctx, tmp_path):
    """dx0 = 20 m traces and dt0 = 2 ms samples: the driver refines the grid to
    10 m and 1.8 ms, so every gather goes through the f-k (lx = 2) and the sinc
    resampling before migration"""
    nz0, nx0, nt0 = 30, 60, 160
    vp0, rho0 = model(nz0, nx0, 23)
    shots = [(200, 0, 12), (460, 40, 14), (700, 20, 10)]          # (sx, offset, traces) at 20 m spacing
    raw, _ = gather_bytes(shots, nt0, dt_us=2000, big=False, sgy=True, seed=24, dstep=20)
    for d in ("input", "output", "seisReal"):
        (tmp_path / d).mkdir()
    (tmp_path / "input" / "cmd.in").write_text("parameter file\npar.txt\n")
    (tmp_path / "input" / "par.txt").write_text(PARFILE.format(nt0=nt0, xmax=(nx0 - 1) * 20, zmax=(nz0 - 1) * 10))
    write_su(str(tmp_path / "input" / "vel.su"), vp0)
    write_su(str(tmp_path / "input" / "rho.su"), rho0)
    (tmp_path / "seisReal" / "gather.sgy").write_bytes(raw)
    assert os.access(EXE, os.X_OK), "rtm_gpu is not built"
    print("[driver-field] rtm_gpu dataflag 1, 3 gathers", flush=True)
    r = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    tab = np.loadtxt(str(tmp_path / "shotprofileId.dat"), dtype=np.int64).T
    assert tab[0].tolist() == [200, 460, 700] and tab[1].tolist() == [12, 14, 10]

    # the same chain through the library's entry points
    g = G.Grid()
    L = G.lib()
    L.rtmh_get_constant(G._fp(vp0), C.c_float(20.0), C.c_float(20.0), C.c_float(10.0), C.c_int(nx0), C.c_int(nz0),
                        C.c_int(nt0), C.c_float(0.002), C.byref(g))
    assert g.lx == 2 and g.lz == 1 and g.nt > nt0
    nx, nz, nt = g.nx, g.nz, g.nt
    vp, rho = np.zeros((nz, nx), np.float32), np.zeros((nz, nx), np.float32)
    L.rtmh_interp_model(G._fp(vp0), G._fp(vp), nx0, nz0, nx, nz, g.lx, g.lz)
    L.rtmh_interp_model(G._fp(rho0), G._fp(rho), nx0, nz0, nx, nz, g.lx, g.lz)
    ltab, e = G.scan_gather(str(tmp_path / "seisReal" / "gather.sgy"), 1)
    assert e == 0 and np.array_equal(ltab, tab)
    fields, recs = [], []
    for k in range(3):
        sx, ntr, itr, off = (int(v) for v in tab[:, k])
        fld, ncol0, ncol = F.gather_apertures(sx, ntr, off, 20.0, g.dx)
        obs = G.read_gather(str(tmp_path / "seisReal" / "gather.sgy"), itr, ntr, nt0, ncol0, sgyflag=1)
        rec = G.interp_seis(obs, ncol, nt, g.dt, 0.002)
        ref = F.interp_seis2d(F.read_gather(raw, itr, ntr, nt0, ncol0, sgyflag=1), ncol, nt, g.dt, 0.002)
        assert np.max(np.abs(rec - ref)) <= 1e-5 * np.max(np.abs(ref))
        fields.append(fld)
        recs.append(rec)
    nw = int(np.float32(1.5) / np.float32(np.float32(20.0) * np.float32(g.dt)))
    tlength = int(np.float32(0.05) / np.float32(g.dt))
    P = G.make_params(nshot=3, startshot=1, dshot=1, mode=1, flag_layout=3, nx=nx, nz=nz, nt=nt, dx=g.dx, dz=g.dz,
                      dt=g.dt, npml=32, fdom=20.0, spz0=20.0, amp=1.0, alp=0.05, tsnap=0.1, sgyflag=1,
                      direct=(nw + 0.5) * 20.0 * g.dt, direct0=(tlength + 0.5) * g.dt)
    lib = gpu_ctx.run(P, vp, rho, seis_in=recs, shot_field=fields, ic_extra=True)
    img1 = read_su(str(tmp_path / "output" / "0Image1.su"), nz, nx)
    ill = read_su(str(tmp_path / "output" / "0Illum.su"), nz, nx)
    assert np.array_equal(img1, lib.image1) and np.array_equal(ill, lib.illum)
    assert np.array_equal(read_su(str(tmp_path / "output" / "0Image2m.su"), nz, nx), lib.image1)
    assert np.abs(img1).max() > 0
