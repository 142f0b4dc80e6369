"""Model smoothing on the GPU (rtmh_smooth_gpu): pmlvelsmooth1d as a
wavefront of lane-shuffled sweeps and velsmooth1d as a per-cell window sum
equal the reference's own outputs (the committed fixture) and the host
restatement (pinned to the reference bit for bit by tests/test_golden.py)
on windows up to the precision configuration's size; a served request with
medium_flag / ismth equals the batch engine's job."""
import ctypes as C

import numpy as np
import pytest

import gubingluo_amd as G
from tests.test_golden import load

pytestmark = pytest.mark.gpu


def host_smooth(w, npml=0, nsp=0):
    a = np.array(w, np.float32, order="C", copy=True)
    nz, nx = a.shape
    if npml:
        G.lib().rtmh_pml_smooth(G._fp(a), C.c_int(nx), C.c_int(nz), C.c_int(npml))
    if nsp:
        G.lib().rtmh_gauss_smooth(G._fp(a), C.c_int(nz), C.c_int(nx), C.c_int(nsp))
    return a


def test_smoothing_golden(gpu_ctx):
    d = load("smoothing.npz")
    base, npml, nsp = d["base"], int(d["npml"]), int(d["nsp"])
    assert np.array_equal(gpu_ctx.smooth(base, npml=npml), d["pml"])
    assert np.array_equal(gpu_ctx.smooth(base, nsp=nsp), d["smooth"])


@pytest.mark.parametrize("nz,nx,npml,nsp,seed", [(60, 90, 20, 3, 1), (124, 225, 32, 5, 2), (612, 401, 50, 10, 3),
                                                 (1124, 2148, 50, 4, 4), (200, 150, 40, 0, 5)])
def test_smoothing_equals_host(gpu_ctx, nz, nx, npml, nsp, seed):
    rng = np.random.default_rng(seed)
    w = rng.uniform(1500.0, 4500.0, (nz, nx)).astype(np.float32)
    w[nz // 3:] += 800.0                       # a contrast the layer relaxation has to spread
    w = w * w                                  # v^2, as the windows hold it
    print(f"[smoothing] {nz}x{nx} npml {npml} nsp {nsp}", flush=True)
    assert np.array_equal(gpu_ctx.smooth(w, npml=npml), host_smooth(w, npml=npml))
    if nsp:
        assert np.array_equal(gpu_ctx.smooth(w, nsp=nsp), host_smooth(w, nsp=nsp))


def test_served_smoothing_equals_engine(gpu_ctx):
    rng = np.random.default_rng(7)
    nz, nx = 60, 160
    vp = np.full((nz, nx), 1900.0, np.float32)
    vp[nz // 2:] = 2600.0
    vp += rng.uniform(-80, 80, (nz, nx)).astype(np.float32)
    rho = (2000.0 + rng.uniform(-100, 100, (nz, nx))).astype(np.float32)
    mid = gpu_ctx.model_create(vp, rho)
    try:
        for mf, sm in ((1, 0), (0, 2), (1, 3)):
            P = G.make_params(nx=nx, nz=nz, nt=300, nshot=3, startshot=1, dshot=1, flag_layout=1, spx0=300.0,
                              dspx=300.0, offsetmax=800.0, npml=32, dx=10.0, dz=10.0, dt=1e-3, fdom=20.0,
                              spz0=20.0, medium_flag=mf, ismth=sm)
            ref = gpu_ctx.run(P, model=mid, batch=2)
            with G.Server(gpu_ctx, max_slots=4) as srv:
                rid = srv.submit(mid, P)
                i1, i2, il, lap, nrm, st = srv.wait(rid, P, timeout=120)
            print(f"[smoothing] served medium_flag {mf} ismth {sm}", flush=True)
            assert np.array_equal(i1, ref.image1) and np.array_equal(i2, ref.image2s)
            assert np.array_equal(il, ref.illum)
    finally:
        gpu_ctx.model_destroy(mid)
