"""Decimated snapshots of the source wavefield (forward pass saves the
interior of p, vx, vz every k steps past the wavelet; the reverse pass
restores them into the reconstructed field): images against the FP64
oracle within 1e-4 (FP32 snapshots) and 1e-3 (FP16 / BF16 snapshots, with
FP16 / BF16 boundary strips too), graph replay equal to plain launches, and
snapshots of a job whose wavefield has left the grid before nt leave the
reconstruction unchanged to rounding.  While the wavefield is still on the
grid at nt the snapshots change the images (they replace the reference's
zero-velocity reconstruction start): there every form equals the
kernel-order oracle with the same store and snapshots bit for bit."""
import numpy as np
import pytest

import gubingluo_amd as G
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def job(nt=1200):
    nz, nx = 64, 200
    rng = np.random.default_rng(21)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[32:] = 2700.0
    vp += rng.uniform(-60, 60, (nz, nx)).astype(np.float32)
    rho = np.full((nz, nx), 2000.0, np.float32)
    kw = dict(nshot=3, startshot=1, dshot=1, flag_layout=1, nx=nx, nz=nz, nt=nt, dx=10.0, dz=10.0, dt=1e-3,
              npml=40, fdom=20.0, spx0=500.0, spz0=20.0, dspx=500.0, offsetmax=800.0)
    return kw, vp, rho


@pytest.fixture(scope="module")
def oracle64():
    kw, vp, rho = job()
    return O.rtm(O.make_params(**kw), vp, rho, f64=True)


@pytest.mark.parametrize("store,snap,tol", [(G.STORE_F32, G.STORE_F32, 1e-4), (G.STORE_F16, G.STORE_F16, 1e-3),
                                            (G.STORE_BF16, G.STORE_BF16, 1e-3)])
def test_snapshots_vs_fp64_oracle(gpu_ctx, oracle64, store, snap, tol):
    kw, vp, rho = job()
    P = G.make_params(**kw)
    print(f"[snapshots] store {store} snapshots {snap} every 100 steps", flush=True)
    r = gpu_ctx.run(P, vp, rho, store_fmt=store, snap_every=100, snap_fmt=snap, batch=3)
    i1, i2, il = oracle64
    e = (rel_l2(r.image1, i1), rel_l2(r.image2s, i2), rel_l2(r.illum, il))
    print(f"[snapshots] image1 {e[0]:.2e} image2s {e[1]:.2e} illum {e[2]:.2e}", flush=True)
    assert max(e) <= tol, e
    # snapshots change the source field only by rounding here (the wavefield
    # has left the grid by nt, so the reconstruction matches the forward field)
    a = gpu_ctx.run(P, vp, rho, store_fmt=store, batch=3)
    assert rel_l2(r.image1, a.image1) <= tol


def test_snapshots_graph_equals_plain(gpu_ctx):
    kw, vp, rho = job(nt=900)
    P = G.make_params(**kw)
    a = gpu_ctx.run(P, vp, rho, store_fmt=G.STORE_F16, snap_every=100, snap_fmt=G.STORE_F16, batch=3,
                    graph_chunk=0, want_shots=True)
    b = gpu_ctx.run(P, vp, rho, store_fmt=G.STORE_F16, snap_every=100, snap_fmt=G.STORE_F16, batch=2,
                    graph_chunk=30, want_shots=True)
    for x, y in zip(a.shot_g1, b.shot_g1):
        assert np.array_equal(x, y)
    assert np.array_equal(a.image1, b.image1) and np.array_equal(a.illum, b.illum)
    # and they do act: a restore replaces the reconstructed interior
    c = gpu_ctx.run(P, vp, rho, store_fmt=G.STORE_F16, batch=3, want_shots=True)
    assert not all(np.array_equal(x, y) for x, y in zip(a.shot_g1, c.shot_g1))


@pytest.mark.parametrize("fmt", [G.STORE_F32, G.STORE_F16, G.STORE_BF16])
def test_snapshots_while_the_wavefield_is_on_the_grid(gpu_ctx, fmt):
    """nt = 500: the wavefield is still on the grid at nt, where the
    reference's zero-velocity start makes its reconstruction depart from the
    forward field and the snapshots change the images.  Each form (FP32, or
    FP16 / BF16 strips and snapshots) equals the kernel-order oracle with the
    same store and snapshots bit for bit, shot by shot; the 16-bit forms are
    also within 1e-3 of the FP64 oracle with exact snapshots."""
    kw, vp, rho = job(nt=500)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    print(f"[snapshots] on the grid, format {fmt}, every 50 steps", flush=True)
    r = gpu_ctx.run(P, vp, rho, store_fmt=fmt, snap_every=50, snap_fmt=fmt, batch=3, want_shots=True)
    for k, ishot in enumerate((1, 2, 3)):
        if fmt == G.STORE_F32:
            o = O.run_shot(Po, ishot, vp, rho, korder=True, snap_every=50)
            assert np.array_equal(r.shot_g1[k], o.g1)
            assert np.array_equal(r.shot_illum[k], o.illum)
        else:
            o = O.run_shot(Po, ishot, vp, rho, korder=True, snap_every=50, store_fmt=fmt, snap_fmt=fmt)
            assert np.array_equal(r.shot_g1[k], o.g1)
            assert np.array_equal(r.shot_illum[k], o.illum)
            o = O.run_shot(Po, ishot, vp, rho, f64=True, snap_every=50)
            e = (rel_l2(r.shot_g1[k], o.g1), rel_l2(r.shot_illum[k], o.illum))
            print(f"[snapshots] shot {ishot} g1 {e[0]:.2e} illum {e[1]:.2e}", flush=True)
            assert max(e) <= 1e-3, e
