"""Linearity and reciprocity of the oracle's modelling and imaging.

Linearity: every update is linear in the fields and the source enters once
per step, so doubling the source amplitude doubles the record and the
receiver wavefield and quadruples the cross-correlation and the
illumination -- exactly in FP64, where nothing reaches the subnormal range
(the record is handed out as float32, whose subnormals are compared apart).
A nonlinear term or an amplitude-independent one (a bias, a floor) breaks it.

Reciprocity: p_t = -K div v + s, rho v_t = -grad p with K = rho v^2 gives
p_tt = K div(rho^-1 grad p) + s_t; K^-1 (d_tt - K div rho^-1 grad) is
self-adjoint, so the pressure record of a pressure source obeys
G(r|s) K(s) = G(s|r) K(r).  The staggered scheme keeps that structure when
its two differences are each other's transposes, 1/rho sits on the velocity
nodes and rho v^2 on the pressure nodes: a shifted stagger, a misplaced
density or a wrong sign would break it.  Checked with a point source on the
surface row (alp = 1/m: exp(-alp^2 dist) is 0 in float one cell away) and a
layered, randomly perturbed subsurface, with density and velocity that
differ at the two surface points."""
import numpy as np
import pytest

from oracle import oracle as O

FLT_MIN = np.float32(1.17549435e-38)


def model(surface_contrast=False):
    nz, nx = 80, 200
    rng = np.random.default_rng(5)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[40:] = 2600.0
    vp[60:] = 3000.0
    vp[10:] += rng.uniform(-100, 100, (nz - 10, nx)).astype(np.float32)
    rho = np.full((nz, nx), 2000.0, np.float32)
    rho[40:] = 2200.0
    if surface_contrast:
        vp[:10, 100:] = 2200.0
        rho[:10, 100:] = 2500.0
    return vp, rho


def shot(vp, rho, spx, amp=1.0, f64=True):
    nz, nx = vp.shape
    kw = dict(nshot=1, startshot=1, dshot=1, flag_layout=2, nx=nx, nz=nz, nt=900, dx=10.0, dz=10.0, dt=1e-3,
              npml=40, fdom=20.0, spx0=spx, spz0=0.0, offsetmax=2000.0, amp=amp, alp=1.0)
    P = O.make_params(**kw)
    return O.run_shot(P, 1, vp, rho, f64=f64), O.geometry(P, 1)


def test_modelling_and_imaging_are_linear_in_the_source():
    vp, rho = model()
    (a, _), (b, _) = shot(vp, rho, 500.0), shot(vp, rho, 500.0, amp=2.0)
    assert np.array_equal(4 * a.g1, b.g1)
    assert np.array_equal(4 * a.illum, b.illum)
    normal = np.abs(a.seis) >= FLT_MIN
    assert normal.sum() > 0.5 * a.seis.size
    assert np.array_equal((2 * a.seis)[normal], b.seis[normal])
    # FP32: the subnormal front of the radius-10 stencil rounds differently,
    # and the difference spreads at the ulp level
    (a32, _), (b32, _) = shot(vp, rho, 500.0, f64=False), shot(vp, rho, 500.0, amp=2.0, f64=False)
    assert np.linalg.norm(2 * a32.seis - b32.seis) <= 1e-6 * np.linalg.norm(b32.seis)


@pytest.mark.parametrize("contrast", [False, True])
def test_source_receiver_reciprocity(contrast):
    vp, rho = model(surface_contrast=contrast)
    (a, ga), (b, gb) = shot(vp, rho, 500.0), shot(vp, rho, 1300.0)
    ta = a.seis[:, gb.nsx - ga.rcol0].astype(np.float64)      # source at x = 500 m, receiver at 1300 m
    tb = b.seis[:, ga.nsx - gb.rcol0].astype(np.float64)      # and the other way round
    # K = rho v^2 at the two surface points (padded row npml is model row 0)
    ka = float(rho[0, 50]) * float(vp[0, 50]) ** 2
    kb = float(rho[0, 130]) * float(vp[0, 130]) ** 2
    assert (ka != kb) == contrast
    assert np.linalg.norm(tb - ta * (ka / kb)) <= 1e-6 * np.linalg.norm(tb)
    # the check resolves a trace: the neighbouring receiver is far off
    tn = a.seis[:, gb.nsx - ga.rcol0 + 1].astype(np.float64)
    assert np.linalg.norm(tn * (ka / kb) - tb) > 0.1 * np.linalg.norm(tb)
