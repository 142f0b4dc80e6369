"""Seeded synthetic velocity models and job parameters for the benchmark
configurations (no field data is available; every input is generated here).

    two_layer_anomaly()  1800/2500 m/s two-layer model with a 3000 m/s Gaussian
                         anomaly (400 x 200, dx=dz=10 m)
    layered_random()     10 dipping layers 1500..4500 m/s plus a Gaussian-
                         filtered 5 % perturbation (sigma 30 m), clipped to
                         [1500, 4725] m/s so the grid guard keeps the grid
"""
from __future__ import annotations

import numpy as np

from . import make_params


def two_layer_anomaly(nx=400, nz=200, dx=10.0, dz=10.0, cx=2000.0, cz=1500.0, radius=200.0):
    x = np.arange(nx, dtype=np.float64) * dx
    z = np.arange(nz, dtype=np.float64) * dz
    Z, X = np.meshgrid(z, x, indexing="ij")
    vb = np.where(Z < 1000.0, 1800.0, 2500.0)
    g = np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / radius ** 2)
    v = vb + (3000.0 - vb) * g
    return v.astype(np.float32)


def layered_random(nx, nz, dx, dz, seed=0, nlayers=10, vmin=1500.0, vmax=4500.0, pert=0.05, sigma_m=30.0,
                   clip=(1500.0, 4725.0)):
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    x = np.arange(nx, dtype=np.float64) * dx
    z = np.arange(nz, dtype=np.float64) * dz
    Z, X = np.meshgrid(z, x, indexing="ij")
    zmax = z[-1]
    depths = np.sort(rng.uniform(0.05 * zmax, 0.95 * zmax, nlayers))
    dips = np.deg2rad(rng.uniform(-10.0, 10.0, nlayers))
    vel_layers = np.linspace(vmin, vmax, nlayers + 1) + rng.uniform(-100.0, 100.0, nlayers + 1)
    v = np.full((nz, nx), vel_layers[0])
    xc = 0.5 * x[-1]
    for k in range(nlayers):
        surf = depths[k] + np.tan(dips[k]) * (X - xc)
        v = np.where(Z >= surf, vel_layers[k + 1], v)
    noise = gaussian_filter(rng.standard_normal((nz, nx)), sigma=(sigma_m / dz, sigma_m / dx))
    noise /= max(noise.std(), 1e-12)
    v = v * (1.0 + pert * noise)
    v = np.clip(v, clip[0], clip[1])
    return v.astype(np.float32)


def single_shot_anomaly():
    """1 shot, 400x200, dx=10, npml 40, Ricker 20 Hz at x=2000 m z=20 m, dt=1 ms nt=2000,
    400 surface receivers (full array), cross-correlation image."""
    vp = two_layer_anomaly()
    rho = np.full_like(vp, 2000.0)
    P = make_params(nshot=1, startshot=1, dshot=1, mode=0, flag_layout=2, nx=400, nz=200, nt=2000,
                    dx=10.0, dz=10.0, dt=1e-3, npml=40, fdom=20.0, spx0=2000.0, spz0=20.0, dspx=50.0,
                    offset=0.0, offsetmax=4000.0, amp=1.0, alp=0.05, direct=1.5, direct0=0.05, tsnap=0.1)
    return P, vp, rho


def survey200(nshot=200, seed=2):
    """200 shots every 50 m, 1024x512, dx=10, npml 50, Ricker 15 Hz, dt=0.8 ms nt=4000,
    split-spread 301 receivers (layout 1, offsetmax 3000 m)."""
    nx, nz, dx = 1024, 512, 10.0
    vp = layered_random(nx, nz, dx, dx, seed=seed)
    rho = np.full_like(vp, 2000.0)
    P = make_params(nshot=nshot, startshot=1, dshot=1, mode=0, flag_layout=1, nx=nx, nz=nz, nt=4000,
                    dx=dx, dz=dx, dt=0.8e-3, npml=50, fdom=15.0, spx0=100.0, spz0=20.0, dspx=50.0,
                    offset=0.0, offsetmax=3000.0, amp=1.0, alp=0.05, direct=1.5, direct0=0.05, tsnap=0.1)
    return P, vp, rho


def batched_ragged(nshot=256, seed=3, window_seed=33):
    """Batched RTM: 2048x768, dx=10, npml 50, Ricker 15 Hz, dt=0.8 ms nt=4000, 256 shots every
    80 m, split-spread windows (layout 1) whose width is drawn per shot from U[400, 1600] columns.
    Returns (P, vp, rho, offsetmax per shot)."""
    nx, nz, dx = 2048, 768, 10.0
    vp = layered_random(nx, nz, dx, dx, seed=seed)
    rho = np.full_like(vp, 2000.0)
    P = make_params(nshot=nshot, startshot=1, dshot=1, mode=0, flag_layout=1, nx=nx, nz=nz, nt=4000,
                    dx=dx, dz=dx, dt=0.8e-3, npml=50, fdom=15.0, spx0=50.0, spz0=20.0, dspx=80.0,
                    offset=0.0, offsetmax=3000.0, amp=1.0, alp=0.05, direct=1.5, direct0=0.05, tsnap=0.1)
    rng = np.random.default_rng(window_seed)
    width = np.floor(rng.uniform(400.0, 1600.0, nshot))
    # layout 1 (rtm_model.cu:124-134): NX = offsetmax/dx + 1 columns, nr = (NX-1)/2 on each side
    offsetmax = ((width - 1.0) * dx).astype(np.float32)
    return P, vp, rho, offsetmax


def precision_grid(nshot=128, seed=4):
    """Precision variant: 2048x1024, dx=dz=7.5 m, npml 50, Ricker 25 Hz, dt=0.6 ms nt=6000,
    128 shots every 120 m, full-aperture receivers (layout 2)."""
    nx, nz, dx = 2048, 1024, 7.5
    vp = layered_random(nx, nz, dx, dx, seed=seed, sigma_m=30.0)
    rho = np.full_like(vp, 2000.0)
    P = make_params(nshot=nshot, startshot=1, dshot=1, mode=0, flag_layout=2, nx=nx, nz=nz, nt=6000,
                    dx=dx, dz=dx, dt=0.6e-3, npml=50, fdom=25.0, spx0=60.0, spz0=15.0, dspx=120.0,
                    offset=0.0, offsetmax=20000.0, amp=1.0, alp=0.05, direct=1.5, direct0=0.05, tsnap=0.1)
    return P, vp, rho


SERVE_NX = (256, 512, 768, 1024, 1280, 1536, 1792, 2048)
SERVE_NZ = (256, 384, 512, 640, 768, 896, 1024)


def serve_models(nmodels=8, seed=5):
    """8 resident models with grids drawn from 256x256 .. 2048x1024 (dx=dz=10 m)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nmodels):
        nx = int(rng.choice(SERVE_NX))
        nz = int(rng.choice(SERVE_NZ))
        vp = layered_random(nx, nz, 10.0, 10.0, seed=100 + k)
        out.append((vp, np.full_like(vp, 2000.0)))
    return out


def serve_stream(nreq, models, rate=4.0, seed=55):
    """Poisson arrivals (rate per second) of RTM requests: model index, nt ~ U[1000, 6000],
    1..32 shots spread evenly over the model, split-spread 3000 m (layout 1), 15 Hz, dt=0.8 ms."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for _ in range(nreq):
        t += rng.exponential(1.0 / rate)
        m = int(rng.integers(0, len(models)))
        nz, nx = models[m][0].shape
        nt = int(rng.integers(1000, 6001))
        ns = int(rng.integers(1, 33))
        step = max(1, nx // ns)
        P = make_params(nshot=ns, startshot=1, dshot=1, mode=0, flag_layout=1, nx=nx, nz=nz, nt=nt,
                        dx=10.0, dz=10.0, dt=0.8e-3, npml=50, fdom=15.0, spx0=10.0 * (step // 2),
                        spz0=20.0, dspx=10.0 * step, offset=0.0, offsetmax=3000.0, amp=1.0, alp=0.05,
                        direct=1.5, direct0=0.05, tsnap=0.1)
        reqs.append((t, m, P))
    return reqs
