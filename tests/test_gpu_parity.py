"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  Bar: seismogram rel-L2 1e-5, image rel-L2 1e-4 (plain
FP32 form) against the reference-order FP32 oracle; the kernels follow the
kernel-order oracle's operations (1/rho, rho*v^2, damping ratios), so the
checks against that form expect bit-identical results."""
import numpy as np
import pytest

import gubingluo_amd as G
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def small_model(nz=60, nx=120, seed=0):
    rng = np.random.default_rng(seed)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[nz // 2:] = 2600.0
    vp += rng.uniform(-50, 50, (nz, nx)).astype(np.float32)
    rho = (2000.0 + rng.uniform(-100, 100, (nz, nx))).astype(np.float32)
    return vp, rho


def job(**kw):
    d = dict(nshot=1, startshot=1, dshot=1, flag_layout=2, nx=120, nz=60, nt=300, dx=10.0, dz=10.0, dt=1e-3,
             npml=32, fdom=20.0, spx0=600.0, spz0=20.0, dspx=50.0, offsetmax=1200.0)
    d.update(kw)
    return d


def to_ref_fields(f, g, npml):
    """full-grid split pair -> the reference's strip arrays"""
    nxp, nzp, nxl = g.nxp, g.nzp, g.nxl
    out = dict(p=f["p"].copy(), vx=np.ascontiguousarray(f["vx"][:, :nxp - 1]),
               vz=np.ascontiguousarray(f["vz"][:nzp - 1, :]))
    for k in ("1", "2"):
        q = f["p" + k]
        out["pl" + k] = np.ascontiguousarray(q[:, :npml])
        out["pr" + k] = np.ascontiguousarray(q[:, nxp - npml:])
        out["pt" + k] = np.ascontiguousarray(q[:npml, npml:nxp - npml])
        out["pb" + k] = np.ascontiguousarray(q[nzp - npml:, npml:nxp - npml])
    return out


@pytest.mark.parametrize("npml,nz,nx", [(32, 60, 120), (40, 60, 120), (32, 170, 300)])
def test_fd_steps_bit_exact(gpu_ctx, npml, nz, nx):
    """forward velocity-stress + PML steps from random fields, incl. source;
    the 300x170 grid also has tiles whose halo is all undamped interior.
    Bit-exact against the kernel-order oracle, 1e-6 against the reference order."""
    vp, rho = small_model(nz=nz, nx=nx)
    kw = job(npml=npml, nz=nz, nx=nx, spx0=10.0 * nx / 2)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    g = G.geometry(P, 1)
    rng = np.random.default_rng(1)
    f = {k: rng.standard_normal((g.nzp, g.nxp)).astype(np.float32) for k in ("p", "vx", "vz", "p1", "p2")}
    f["vx"][:, -1] = 0
    f["vz"][-1, :] = 0
    # consistent split pair in the PML (p = p1 + p2 there)
    ref = to_ref_fields(f, g, npml)
    ref32 = {k: v.copy() for k, v in ref.items()}
    print(f"[fd_steps] npml={npml} nxp={g.nxp} nzp={g.nzp}", flush=True)
    gpu = {k: v.copy() for k, v in f.items()}
    gpu_ctx.fd_steps(P, 1, vp, rho, gpu, 7, it0=40, with_source=True)
    O.fd_steps(Po, 1, vp, rho, ref, 7, it0=40, with_source=True, korder=True)
    O.fd_steps(Po, 1, vp, rho, ref32, 7, it0=40, with_source=True)
    for k in ("p", "vx", "vz"):
        assert rel_l2(ref[k], ref32[k]) <= 1e-6, k
    assert np.array_equal(gpu["p"], ref["p"]), rel_l2(gpu["p"], ref["p"])
    assert np.array_equal(gpu["vx"][:, :g.nxp - 1], ref["vx"])
    assert np.array_equal(gpu["vz"][:g.nzp - 1], ref["vz"])
    back = to_ref_fields(gpu, g, npml)
    for k in ("pl1", "pl2", "pr1", "pr2", "pt1", "pt2", "pb1", "pb2"):
        assert np.array_equal(back[k], ref[k]), k


@pytest.mark.parametrize("layout,npml", [(2, 32), (1, 32), (2, 40), (3, 40), (4, 40)])
def test_single_shot_vs_oracle(gpu_ctx, layout, npml):
    vp, rho = small_model(seed=layout)
    kw = job(flag_layout=layout, npml=npml, offset=100.0, offsetmax=600.0)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    print(f"[single_shot] layout={layout} npml={npml}", flush=True)
    r = gpu_ctx.run(P, vp, rho, want_shots=True)
    o = O.run_shot(Po, 1, vp, rho)
    k = O.run_shot(Po, 1, vp, rho, korder=True)
    assert rel_l2(r.shot_seis[0], o.seis) <= 1e-5
    assert rel_l2(r.shot_g1[0], o.g1) <= 1e-4
    assert rel_l2(r.shot_illum[0], o.illum) <= 1e-4
    assert np.array_equal(r.shot_seis[0], k.seis)
    assert np.array_equal(r.shot_g1[0], k.g1)
    assert np.array_equal(r.shot_illum[0], k.illum)


def test_multi_shot_images_vs_oracle(gpu_ctx):
    vp, rho = small_model(nz=50, nx=160, seed=3)
    kw = job(nx=160, nz=50, nshot=5, dshot=2, flag_layout=1, offsetmax=600.0, spx0=300.0, dspx=120.0, nt=250)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    print("[multi_shot] 3 shots layout 1", flush=True)
    r = gpu_ctx.run(P, vp, rho, batch=2)
    o1, o2, ol = O.rtm(Po, vp, rho, stack64=True)
    assert rel_l2(r.image1, o1) <= 1e-4
    assert rel_l2(r.image2s, o2) <= 1e-4
    assert rel_l2(r.illum, ol) <= 1e-4
    i1, i2, il = O.rtm(Po, vp, rho, stack64=True, korder=True)
    assert np.array_equal(r.image1, i1) and np.array_equal(r.image2s, i2) and np.array_equal(r.illum, il)
    assert np.array_equal(r.image_lap, O.laplacian(i1, P.dx, P.dz))
    assert np.array_equal(r.image_norm, O.normalize_multi(i1, il))


def test_batch_and_graph_invariance(gpu_ctx):
    vp, rho = small_model(nz=50, nx=160, seed=4)
    kw = job(nx=160, nz=50, nshot=4, flag_layout=1, offsetmax=600.0, spx0=300.0, dspx=150.0, nt=230)
    P = G.make_params(**kw)
    print("[invariance] batch 1 vs 4, graph chunk 0 vs 16", flush=True)
    a = gpu_ctx.run(P, vp, rho, batch=1, graph_chunk=0, want_shots=True)
    b = gpu_ctx.run(P, vp, rho, batch=4, graph_chunk=16, want_shots=True)
    for x, y in zip(a.shot_g1, b.shot_g1):
        assert np.array_equal(x, y)
    assert np.array_equal(a.image1, b.image1) and np.array_equal(a.image2s, b.image2s)


def test_mode1_migrates_given_data(gpu_ctx):
    vp, rho = small_model(seed=5)
    kw = job()
    P = G.make_params(**kw)
    print("[mode1]", flush=True)
    r0 = gpu_ctx.run(P, vp, rho, want_shots=True)
    P1 = G.make_params(**dict(kw, mode=1))
    r1 = gpu_ctx.run(P1, vp, rho, want_shots=True, seis_in=r0.shot_seis)
    assert np.array_equal(r0.image1, r1.image1)


def test_large_grid_single_shot_vs_oracle(gpu_ctx):
    """grid large enough for interior tiles (undamped halo: the fast path)"""
    vp, rho = small_model(nz=170, nx=300, seed=7)
    kw = job(nx=300, nz=170, nt=260, flag_layout=1, offsetmax=1200.0, spx0=1500.0, npml=32)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    print("[large_grid] 300x170 layout 1", flush=True)
    r = gpu_ctx.run(P, vp, rho, want_shots=True, tile_z=32)
    o32 = O.run_shot(Po, 1, vp, rho)
    assert rel_l2(r.shot_seis[0], o32.seis) <= 1e-5 and rel_l2(r.shot_g1[0], o32.g1) <= 1e-4
    o = O.run_shot(Po, 1, vp, rho, korder=True)
    assert np.array_equal(r.shot_seis[0], o.seis)
    assert np.array_equal(r.shot_g1[0], o.g1)
    assert np.array_equal(r.shot_illum[0], o.illum)


def test_tile_heights_and_autotuner_invariant(gpu_ctx):
    """every tile height (and the autotuned choice) computes the same cells"""
    vp, rho = small_model(nz=150, nx=260, seed=8)
    kw = job(nx=260, nz=150, nshot=3, flag_layout=1, offsetmax=900.0, spx0=700.0, dspx=600.0, nt=240)
    P = G.make_params(**kw)
    print("[tile_heights] 16/32/64/auto", flush=True)
    runs = [gpu_ctx.run(P, vp, rho, batch=3, tile_z=tz, want_shots=True) for tz in (16, 32, 64, 0)]
    assert runs[3].stats.tile_z in (16, 32, 64)
    for r in runs[1:]:
        for x, y in zip(runs[0].shot_g1, r.shot_g1):
            assert np.array_equal(x, y)
        for x, y in zip(runs[0].shot_seis, r.shot_seis):
            assert np.array_equal(x, y)
        assert np.array_equal(runs[0].image1, r.image1) and np.array_equal(runs[0].illum, r.illum)


def test_ragged_windows_vs_oracle(gpu_ctx):
    """one batch of shots with different aperture windows (per-shot offsetmax)"""
    vp, rho = small_model(nz=60, nx=260, seed=9)
    kw = job(nx=260, nz=60, nshot=4, flag_layout=1, spx0=400.0, dspx=500.0, nt=220)
    om = np.array([300.0, 900.0, 550.0, 1200.0], np.float32)
    P = G.make_params(**kw)
    print("[ragged] per-shot offsetmax", om.tolist(), flush=True)
    r = gpu_ctx.run(P, vp, rho, batch=4, want_shots=True, shot_offsetmax=om)
    for k in range(4):
        Po = O.make_params(**dict(kw, offsetmax=float(om[k]), startshot=k + 1, nshot=k + 1))
        o = O.run_shot(Po, k + 1, vp, rho, korder=True)
        assert r.shot_g1[k].shape == o.g1.shape
        assert np.array_equal(r.shot_seis[k], o.seis)
        assert np.array_equal(r.shot_g1[k], o.g1)


@pytest.mark.parametrize("medium_flag,ismth", [(1, 0), (0, 2), (1, 3)])
def test_model_smoothing_vs_oracle(gpu_ctx, medium_flag, ismth):
    """medium_flag: the absorbing layer of each shot's window relaxed before
    the forward pass (pmlvelsmooth1d, also changing vpmax and the damping);
    ismth: the window Gaussian-smoothed before the reverse pass (velsmooth1d).
    Two shots with different windows in one batch."""
    vp, rho = small_model(nz=60, nx=200, seed=12)
    kw = job(nx=200, nz=60, nshot=2, dshot=1, flag_layout=1, npml=32, spx0=700.0, dspx=500.0, offsetmax=800.0,
             nt=260, medium_flag=medium_flag, ismth=ismth)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    print(f"[smoothing] medium_flag={medium_flag} ismth={ismth}", flush=True)
    r = gpu_ctx.run(P, vp, rho, batch=2, want_shots=True)
    plain = gpu_ctx.run(G.make_params(**dict(kw, medium_flag=0, ismth=0)), vp, rho, batch=2, want_shots=True)
    for k in range(2):
        o = O.run_shot(Po, k + 1, vp, rho, korder=True)
        assert np.array_equal(r.shot_seis[k], o.seis), k
        assert np.array_equal(r.shot_g1[k], o.g1), k
        assert np.array_equal(r.shot_illum[k], o.illum), k
        o32 = O.run_shot(Po, k + 1, vp, rho)
        assert rel_l2(r.shot_seis[k], o32.seis) <= 1e-5 and rel_l2(r.shot_g1[k], o32.g1) <= 1e-4
        assert not np.array_equal(r.shot_g1[k], plain.shot_g1[k])      # the smoothing took effect


def test_checkpoint_resume_equals_uninterrupted(gpu_ctx, tmp_path):
    """a job stopped after 3 of 5 shots and resumed from its checkpoint gives the
    uninterrupted job's images bit for bit; another job's checkpoint is refused"""
    vp, rho = small_model(nz=50, nx=160, seed=13)
    kw = job(nx=160, nz=50, nshot=5, dshot=1, flag_layout=1, offsetmax=600.0, spx0=300.0, dspx=120.0, nt=220)
    print("[checkpoint] 5 shots vs 3 + resume", flush=True)
    full = gpu_ctx.run(G.make_params(**kw), vp, rho, batch=2)
    ck = str(tmp_path / "job.ckpt")
    first = gpu_ctx.run(G.make_params(**dict(kw, nshot=3)), vp, rho, batch=2, checkpoint=ck)
    assert first.stats.batches == 2
    rest = gpu_ctx.run(G.make_params(**kw), vp, rho, batch=2, checkpoint=ck, resume=True)
    assert rest.stats.batches == 1                           # shots 4 and 5 only
    for name in ("image1", "image2s", "illum", "image_lap", "image_norm"):
        assert np.array_equal(getattr(rest, name), getattr(full, name)), name
    with pytest.raises(RuntimeError, match="another job"):
        gpu_ctx.run(G.make_params(**dict(kw, fdom=25.0)), vp, rho, batch=2, checkpoint=ck, resume=True)


def test_other_imaging_conditions_vs_oracle(gpu_ctx):
    """decomposition, Poynting and energy-norm images (ic_extra): the kernel-order
    oracle's per-shot products stacked as the engine stacks them (taper, FP64,
    illumination-normalised or not).  Decomposition and energy norm bit for bit;
    the Poynting images within 1e-3 (acosf / cosf / sinf of the device library
    against the host's)"""
    nz, nx = 50, 160
    vp, rho = small_model(nz=nz, nx=nx, seed=14)
    kw = job(nx=nx, nz=nz, nshot=5, dshot=2, flag_layout=1, offsetmax=600.0, spx0=300.0, dspx=120.0, nt=230)
    P, Po = G.make_params(**kw), O.make_params(**kw)
    print("[imaging conditions] 3 shots layout 1", flush=True)
    r = gpu_ctx.run(P, vp, rho, batch=2, ic_extra=True)
    base = gpu_ctx.run(P, vp, rho, batch=2)
    assert np.array_equal(r.image1, base.image1) and np.array_equal(r.illum, base.illum)
    st = {k: np.zeros((nz, nx)) for k in G.IC_IMAGES}
    nwin = 50
    for ishot in (1, 3, 5):
        o = O.run_shot(Po, ishot, vp, rho, korder=True, extra_ic=True)
        g, nxl = o.geom, o.geom.nxl
        while 2 * nwin >= nxl:
            nwin -= 3
        t = O.taper_row(nxl, nwin)
        il = (o.illum * t).astype(np.float64)
        cols = slice(g.nx1, g.nx1 + nxl)
        for k, name in (("3", "g2"), ("4", "g3"), ("5", "g4")):
            gk = (o.ic[name] * t).astype(np.float64)
            st[f"image{k}s"][:, cols] += gk / (il + 1e-6)
            st[f"image{k}m"][:, cols] += gk
        st["imagetrue"][:, cols] += (o.ic["g3_true"] * t).astype(np.float64)
    for k in ("image3s", "image3m", "image5s", "image5m"):
        assert np.array_equal(r.ic[k], st[k].astype(np.float32)), (k, rel_l2(r.ic[k], st[k]))
    for k in ("image4s", "image4m", "imagetrue"):
        assert rel_l2(r.ic[k], st[k]) <= 1e-3, (k, rel_l2(r.ic[k], st[k]))
    for k in G.IC_IMAGES:
        assert np.abs(r.ic[k]).max() > 0, k
