"""Snapshot semantics of the oracle (the library's snap_every option, not the
reference's semantics), pinned by what the leapfrog fixes.

The reference starts its reverse reconstruction from the last pressure with
the source velocities set to zero (rtm_model.cu:366-372), so while a
wavefield is still on the grid at nt its reconstruction departs from the
forward field.  A snapshot of step k holds p of forward step k and the
velocities of forward step k+1 -- the state the reverse leapfrog holds after
reverse step k -- so from a restore on, the reconstruction must reproduce the
forward field to rounding; a snapshot pairing p with the velocities of the
same step would leave an O(dt) error (several per cent here), as would the
reference's zero-velocity start without snapshots."""
import numpy as np
import pytest

from oracle import oracle as O


def job(nt=500):
    nz, nx = 64, 200
    rng = np.random.default_rng(21)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[32:] = 2700.0
    vp += rng.uniform(-60, 60, (nz, nx)).astype(np.float32)
    rho = np.full((nz, nx), 2000.0, np.float32)
    kw = dict(nshot=3, startshot=1, dshot=1, flag_layout=1, nx=nx, nz=nz, nt=nt, dx=10.0, dz=10.0, dt=1e-3,
              npml=40, fdom=20.0, spx0=500.0, spz0=20.0, dspx=500.0, offsetmax=800.0)
    return O.make_params(**kw), vp, rho


def rel(a, b):
    return float(np.linalg.norm((a - b).astype(np.float64)) / np.linalg.norm(a.astype(np.float64)))


@pytest.mark.parametrize("f64", [True, False])
def test_reconstruction_between_snapshots_equals_forward(f64):
    # t0 = ceil(4 / (20 Hz * 1 ms)) = 200, every 50: snapshots at steps 200..450
    P, vp, rho = job()
    n = P.npml
    tol = 1e-12 if f64 else 1e-6
    for it in (449, 420, 401, 330, 251):
        r = O.run_shot(P, 2, vp, rho, snap_it=it, f64=f64, snap_every=50)
        e = rel(r.snap_fwd[n:-n, n:-n], r.snap_rec[n:-n, n:-n])
        assert e <= tol, (it, e)
        # without snapshots the zero-velocity start leaves a large error
        r0 = O.run_shot(P, 2, vp, rho, snap_it=it, f64=f64)
        assert rel(r0.snap_fwd[n:-n, n:-n], r0.snap_rec[n:-n, n:-n]) > 1e-2


def test_snapshot_plan_starts_past_the_wavelet():
    # before the first snapshot (step 200) in reverse time the reconstruction
    # is the reference's again: the restores change nothing below k0 * every
    P, vp, rho = job()
    n = P.npml
    a = O.run_shot(P, 2, vp, rho, snap_it=450, snap_every=50)
    b = O.run_shot(P, 2, vp, rho, snap_it=450)
    # step 450 is the first restore in reverse order: its pre-restore field is
    # the reference's reconstruction
    assert np.array_equal(a.snap_rec, b.snap_rec)
    assert np.array_equal(a.snap_fwd, b.snap_fwd)
    # the forward pass and the record are untouched by snapshots
    assert np.array_equal(a.seis, b.seis)


def test_snapshots_after_the_wave_has_left_change_only_rounding():
    # nt = 1200: the wavefield has left the grid long before nt, so the
    # reference's reconstruction is already the forward field and restoring
    # it changes the images by rounding only
    P, vp, rho = job(nt=1200)
    a = O.run_shot(P, 2, vp, rho, f64=True, snap_every=100)
    b = O.run_shot(P, 2, vp, rho, f64=True)
    assert rel(b.g1, a.g1) <= 1e-4
    assert rel(b.illum, a.illum) <= 1e-4
