"""Reduced-precision boundary store (FP16 / BF16 with per-step max-abs
scaling, FP32 stencil, FP64 stacks) against the FP64 oracle: the images stay
within 1e-3 relative L2; the FP32 form within 1e-4.  Against the kernel-order
oracle with the same 16-bit store (orc_store16_block): bit for bit."""
import numpy as np
import pytest

import gubingluo_amd as G
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def job():
    nz, nx = 64, 200
    rng = np.random.default_rng(11)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[30:] = 2700.0
    vp += rng.uniform(-60, 60, (nz, nx)).astype(np.float32)
    rho = np.full((nz, nx), 2000.0, np.float32)
    kw = dict(nshot=5, startshot=1, dshot=2, flag_layout=1, nx=nx, nz=nz, nt=500, dx=10.0, dz=10.0, dt=1e-3,
              npml=40, fdom=20.0, spx0=500.0, spz0=20.0, dspx=200.0, offsetmax=800.0)
    return kw, vp, rho


@pytest.fixture(scope="module")
def oracle64():
    kw, vp, rho = job()
    return O.rtm(O.make_params(**kw), vp, rho, f64=True)


@pytest.mark.parametrize("fmt,tol", [(G.STORE_F32, 1e-4), (G.STORE_F16, 1e-3), (G.STORE_BF16, 1e-3)])
def test_store_precision_vs_fp64_oracle(gpu_ctx, oracle64, fmt, tol):
    kw, vp, rho = job()
    print(f"[precision] store_fmt={fmt}", flush=True)
    r = gpu_ctx.run(G.make_params(**kw), vp, rho, store_fmt=fmt, batch=3)
    i1, i2, il = oracle64
    e1, e2, e3 = rel_l2(r.image1, i1), rel_l2(r.image2s, i2), rel_l2(r.illum, il)
    print(f"[precision] fmt={fmt} image1 {e1:.2e} image2s {e2:.2e} illum {e3:.2e}", flush=True)
    assert e1 <= tol and e2 <= tol and e3 <= tol


def test_store_16bit_graph_matches_plain(gpu_ctx):
    kw, vp, rho = job()
    P = G.make_params(**kw)
    a = gpu_ctx.run(P, vp, rho, store_fmt=G.STORE_F16, batch=3, graph_chunk=0)
    b = gpu_ctx.run(P, vp, rho, store_fmt=G.STORE_F16, batch=1, graph_chunk=32)
    assert np.array_equal(a.image1, b.image1) and np.array_equal(a.image2s, b.image2s)


@pytest.mark.parametrize("fmt", [G.STORE_F16, G.STORE_BF16])
def test_store_16bit_bit_exact_vs_oracle(gpu_ctx, fmt):
    kw, vp, rho = job()
    print(f"[precision] store_fmt={fmt} against the kernel-order oracle", flush=True)
    r = gpu_ctx.run(G.make_params(**kw), vp, rho, store_fmt=fmt, batch=3)
    i1, i2, il = O.rtm(O.make_params(**kw), vp, rho, stack64=True, korder=True, store_fmt=fmt)
    for name, a, b in (("image1", r.image1, i1), ("image2s", r.image2s, i2), ("illum", r.illum, il)):
        assert np.array_equal(a, b), (name, rel_l2(a, b))


@pytest.mark.parametrize("fmt", [G.STORE_F16, G.STORE_BF16])
def test_store_16bit_wide_window_bit_exact_vs_oracle(gpu_ctx, fmt):
    """a 2600-column window: its top/bottom strip blocks hold 2 x 10 x 2600 =
    52000 values, more than the save kernel keeps in registers (48 x 1024), so
    the block is read twice -- still bit for bit the oracle's store"""
    nz, nx = 40, 2600
    rng = np.random.default_rng(12)
    vp = (2200.0 + rng.uniform(-80, 80, (nz, nx))).astype(np.float32)
    rho = np.full((nz, nx), 2000.0, np.float32)
    kw = dict(nshot=1, startshot=1, dshot=1, flag_layout=2, nx=nx, nz=nz, nt=150, dx=10.0, dz=10.0, dt=1e-3,
              npml=20, fdom=20.0, spx0=13000.0, spz0=20.0, offsetmax=13000.0)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    assert 20 * G.geometry(P, 1).nxl > 48 * 1024
    print(f"[precision] wide window, store_fmt={fmt}", flush=True)
    r = gpu_ctx.run(P, vp, rho, store_fmt=fmt, want_shots=True)
    o = O.run_shot(Po, 1, vp, rho, korder=True, store_fmt=fmt)
    assert np.array_equal(r.shot_g1[0], o.g1) and np.array_equal(r.shot_illum[0], o.illum)
