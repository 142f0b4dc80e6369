"""The drop-in driver end to end on the GPU: the `rtm_gpu` executable (the
reference's RTM_GPU_velstr.cpp role: cmd.in, parameter file, SU velocity and
density, selfdoc, rtm_model, SU image outputs) run as a child process in a job
directory, its images read back from the .su files and compared bit for bit
with the library's own entry point and with the kernel-order oracle."""
import os
import subprocess

import numpy as np
import pytest

import gubingluo_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gubingluo_amd", "rtm_gpu")

PARFILE = """dt(s) nt tsnap(s)
0.001 {nt} 0.1
xmax zmax dx dz (m)
{xmax} {zmax} 10 10
fdom amp alp
20 1 0.05
spx0 spz0 dspx
300 20 120
offset offsetmax
0 600
startshot nshot dshot nshotall
1 {nshot} 2 {nshot}
npml
32
medium_flag ismth
0 0
direct direct0
1.5 0.05
idevice
0
flag_layout
1
mode
0
dataflag
0
sgyflag
0
endianflag
0
rhoflag
1
vp_file
vel.su
rho_file
rho.su
"""


def write_su(path, a):
    """nx traces of nz float32 samples, 240-byte headers (the layout the
    reference's readvelocity skips over)"""
    nz, nx = a.shape
    rec = np.zeros(nx, dtype=[("h", "V240"), ("d", "<f4", (nz,))])
    rec["d"] = np.ascontiguousarray(a.T, np.float32)
    rec.tofile(path)


def read_su(path, nz, nx):
    rec = np.fromfile(path, dtype=[("h", "V240"), ("d", "<f4", (nz,))])
    assert rec.shape == (nx,)
    return np.ascontiguousarray(rec["d"].T)


def test_rtm_gpu_driver_images(gpu_ctx, tmp_path):
    nz, nx, nt, nshot = 50, 160, 250, 5
    rng = np.random.default_rng(11)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[nz // 2:] = 2600.0
    vp += rng.uniform(-50, 50, (nz, nx)).astype(np.float32)
    rho = (2000.0 + rng.uniform(-100, 100, (nz, nx))).astype(np.float32)
    for d in ("input", "output", "seisReal"):
        (tmp_path / d).mkdir()
    (tmp_path / "input" / "cmd.in").write_text("parameter file\npar.txt\n")
    (tmp_path / "input" / "par.txt").write_text(
        PARFILE.format(nt=nt, xmax=(nx - 1) * 10, zmax=(nz - 1) * 10, nshot=nshot))
    write_su(str(tmp_path / "input" / "vel.su"), vp)
    write_su(str(tmp_path / "input" / "rho.su"), rho)
    assert os.access(EXE, os.X_OK), "rtm_gpu is not built"
    print("[driver] rtm_gpu 160x50 nt 250, shots 1,3,5", flush=True)
    r = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Prestack Reverse Time Migration by Bingluo Gu from UPC." in r.stdout
    out = {n: read_su(str(tmp_path / "output" / f"0{n}.su"), nz, nx)
           for n in ("Image1", "Image2s", "Illum", "Image2m", "Image3s", "Image4s", "Image5m", "Image1lap", "Imagenorm")}
    true = read_su(str(tmp_path / "output" / "0Imagetrue.su"), nz, nx)
    for k in (1, 3, 5):
        assert (tmp_path / "seisReal" / f"{k}seisreal.su").stat().st_size > 0
    # the same job through the library's own entry point and through the oracle
    kw = dict(nshot=nshot, startshot=1, dshot=2, flag_layout=1, nx=nx, nz=nz, nt=nt, dx=10.0, dz=10.0, dt=1e-3,
              npml=32, fdom=20.0, spx0=300.0, spz0=20.0, dspx=120.0, offset=0.0, offsetmax=600.0)
    lib = gpu_ctx.run(G.make_params(**kw), vp, rho, ic_extra=True)   # the drop-in computes every image
    assert np.array_equal(out["Image1"], lib.image1)
    assert np.array_equal(out["Image2s"], lib.image2s)
    assert np.array_equal(out["Illum"], lib.illum)
    assert np.array_equal(out["Image2m"], lib.image1)          # image2m = image1 in the reference
    assert np.array_equal(out["Image3s"], lib.ic["image3s"]) and out["Image3s"].any()
    assert np.array_equal(out["Image4s"], lib.ic["image4s"]) and np.array_equal(out["Image5m"], lib.ic["image5m"])
    assert np.array_equal(true, lib.ic["imagetrue"])
    assert np.array_equal(out["Image1lap"], lib.image_lap)
    assert np.array_equal(out["Imagenorm"], lib.image_norm)
    i1, i2, il = O.rtm(O.make_params(**kw), vp, rho, stack64=True, korder=True)
    assert np.array_equal(out["Image1"], i1) and np.array_equal(out["Image2s"], i2) and np.array_equal(out["Illum"], il)
    assert np.abs(i1).max() > 0
    # RTMH_CHECKPOINT: the first run writes the checkpoint, the second finds every
    # shot done and writes the same images from the stacks it read back
    env = dict(os.environ, RTMH_CHECKPOINT=str(tmp_path / "job.ckpt"))
    for rep in range(2):
        r = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=150, env=env)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert (tmp_path / "job.ckpt").stat().st_size > 3 * nx * nz * 8
        for n in ("Image1", "Image2s", "Illum"):
            assert np.array_equal(read_su(str(tmp_path / "output" / f"0{n}.su"), nz, nx), out[n]), (rep, n)
    assert "shot finished" not in r.stdout                          # the second run ran no shot
