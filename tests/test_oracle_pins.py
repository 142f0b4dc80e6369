"""Pins of the CPU oracle (runs on the CPU, no GPU).

The reference has no tests or fixtures, so each part of the oracle is pinned
by something the mathematics fixes or by an invariant a plausible mistake
would break (sign, staggering direction, index of a coefficient row, a
transposed operand, left/right or top/bottom PML ranges):

  coefficients   consistency sum (2i-1)c_i = 1 and the published sum |c|
  velocity step  one step from p = sin(kx) gives vx = -dt/rho * dp/dx at x+dx/2
  pressure step  one step from vx = sin(k(x+dx/2)) gives p = -dt rho v^2 dvx/dx
  modelling      homogeneous medium vs the analytic 2-D Green's function
  symmetry       mirror-symmetric model and source -> exactly mirrored records
  PML            energy left on the grid after the wave leaves is small
  literal form   literal kernel ranges == generalised form, bit for bit, npml 32
  reconstruction boundary-saved reverse propagation reproduces the forward field
  precision      FP32 restatement vs FP64 within the parity tolerances
  wavelet/taper/mute/Laplacian closed forms
"""
import numpy as np
import pytest
from scipy.special import hankel1

from oracle import oracle as O
from tests.conftest import rel_l2


def test_coefficient_consistency():
    T = O.coef_table().reshape(10, 10)
    for r in range(1, 11):
        c = T[r - 1, :r].astype(np.float64)
        assert abs(np.sum((2 * np.arange(1, r + 1) - 1) * c) - 1.0) < 2e-6, r
        assert np.all(T[r - 1, r:] == 0)
    assert abs(np.abs(T[9]).astype(np.float64).sum() - 1.4760199) < 1e-6


def _homog(nz, nx, v=2000.0, rho=2000.0):
    return np.full((nz, nx), v, np.float32), np.full((nz, nx), rho, np.float32)


def _fields(P):
    g = O.geometry(P, 1)
    return g, O.alloc_fields(g, P.npml, f64=True)


def test_velocity_step_is_staggered_derivative():
    nz, nx = 40, 80
    vp, rho = _homog(nz, nx, 2500.0, 1800.0)
    P = O.make_params(nx=nx, nz=nz, npml=16, nt=10, dt=1e-3, dx=10.0, dz=12.0)
    g, F = _fields(P)
    k = 2 * np.pi / (12 * P.dx)
    x = np.arange(g.nxp) * P.dx
    z = np.arange(g.nzp) * P.dz
    F["p"][:] = np.sin(k * x)[None, :]
    O.fd_steps(P, 1, vp, rho, F, 1, f64=True)
    xs = (np.arange(g.nxp - 1) + 0.5) * P.dx
    want = -P.dt / 1800.0 * k * np.cos(k * xs)
    sl = (slice(30, -30), slice(30, -30))           # away from the absorbing layer
    # the optimised 20th-order stencil is accurate to ~2e-5 at kh = 0.52; a wrong
    # stagger, sign or coefficient row is wrong by O(1)
    assert rel_l2(F["vx"][sl], np.broadcast_to(want, F["vx"].shape)[sl]) < 2e-4
    # z direction
    g, F = _fields(P)
    F["p"][:] = np.sin(k * z)[:, None]
    O.fd_steps(P, 1, vp, rho, F, 1, f64=True)
    zs = (np.arange(g.nzp - 1) + 0.5) * P.dz
    want = -P.dt / 1800.0 * k * np.cos(k * zs)
    assert rel_l2(F["vz"][sl], np.broadcast_to(want[:, None], F["vz"].shape)[sl]) < 2e-4


def test_pressure_step_is_staggered_divergence():
    nz, nx = 40, 80
    vp, rho = _homog(nz, nx, 2500.0, 1800.0)
    P = O.make_params(nx=nx, nz=nz, npml=16, nt=10, dt=1e-3, dx=10.0, dz=12.0)
    g, F = _fields(P)
    k = 2 * np.pi / (10 * P.dx)
    xs = (np.arange(g.nxp - 1) + 0.5) * P.dx
    F["vx"][:] = np.sin(k * xs)[None, :]
    O.fd_steps(P, 1, vp, rho, F, 1, f64=True)      # p = 0, so the velocity step leaves vx unchanged
    x = np.arange(g.nxp) * P.dx
    want = -P.dt * 1800.0 * 2500.0 ** 2 * k * np.cos(k * x)
    sl = (slice(30, -30), slice(30, -30))
    assert rel_l2(F["p"][sl], np.broadcast_to(want, F["p"].shape)[sl]) < 2e-4


def test_homogeneous_seismogram_matches_green_function():
    """p_t = -rho v^2 div v + s(x) w(t), v_t = -grad p / rho gives
    p_tt - v^2 lap p = s(x) w'(t): p = S * (w' conv u), u the 2-D Green's
    function, S = amp*pi/alp^2 the strength of the Gaussian source.  The
    discrete source enters after the p-update of step it with w[it] (half a
    step ahead of the continuous time, taken into account).  Outside its
    support the radial Gaussian radiates as a point source of spectrum
    S exp(-k^2/(4 alp^2)), k = omega/v (Hankel addition theorem)."""
    nz, nx, v = 160, 200, 2000.0
    vp, rho = _homog(nz, nx, v)
    nt, dt, f = 1200, 5e-4, 15.0
    P = O.make_params(nx=nx, nz=nz, npml=40, nt=nt, dt=dt, fdom=f, spx0=1000.0, spz0=200.0, flag_layout=2,
                      alp=0.05, amp=1.0)
    r = O.run_shot(P, 1, vp, rho, forward_only=True, f64=True)
    g = r.geom
    ir = 140                                  # receiver 400 m off, on the surface row
    dist = np.hypot((ir + g.rcol0 - g.nsx) * P.dx, (P.npml - g.nsz) * P.dz)
    assert dist / v + 2.0 / f < nt * dt       # the direct wave is inside the record
    sim = r.seis[:, ir].astype(np.float64)
    n = 16384
    t = (np.arange(n) + 0.5) * dt
    tau = np.pi * f * (t - 1.0 / f)
    wv = (1 - 2 * tau ** 2) * np.exp(-tau ** 2)
    W = np.fft.rfft(wv)                       # numpy: X(om) = sum x exp(-i om t)
    om = 2 * np.pi * np.fft.rfftfreq(n, dt)
    U = np.zeros_like(W)
    # lap U + (om/v)^2 U = -delta / v^2, outgoing for the exp(+i om t) synthesis
    U[1:] = np.conj(1j / 4.0 * hankel1(0, om[1:] * dist / v)) / v ** 2
    S = np.pi / P.alp ** 2 * P.amp * np.exp(-(om / v) ** 2 / (4.0 * P.alp ** 2))
    ana = np.fft.irfft(S * (1j * om) * W * U, n)[:nt]
    corr = np.dot(sim, ana) / np.linalg.norm(sim) / np.linalg.norm(ana)
    amp_ratio = np.linalg.norm(sim) / np.linalg.norm(ana)
    assert corr > 0.999, corr
    assert abs(amp_ratio - 1.0) < 0.01, amp_ratio


def test_mirror_symmetry_is_exact():
    nz, nx = 50, 101
    rng = np.random.default_rng(0)
    half = rng.uniform(1800, 2600, (nz, 51)).astype(np.float32)
    vp = np.concatenate([half, half[:, :50][:, ::-1]], axis=1)
    rho = np.full_like(vp, 2000.0)
    P = O.make_params(nx=nx, nz=nz, npml=20, nt=250, spx0=500.0, spz0=30.0, flag_layout=2)
    r = O.run_shot(P, 1, vp, rho, forward_only=True)
    assert r.geom.nsx == r.geom.nxp // 2
    assert np.array_equal(r.seis, r.seis[:, ::-1])


def test_pml_absorbs():
    nz, nx = 80, 80
    vp, rho = _homog(nz, nx, 2000.0)
    P = O.make_params(nx=nx, nz=nz, npml=30, nt=1500, dt=1e-3, spx0=400.0, spz0=400.0, flag_layout=2)
    e = []
    for it in (250, 1499):
        r = O.run_shot(P, 1, vp, rho, forward_only=True, snap_it=it, f64=True)
        s = r.snap_fwd[30:-30, 30:-30]
        e.append(float(np.sum(s.astype(np.float64) ** 2)))
    assert e[1] < 1e-4 * e[0], e


@pytest.mark.parametrize("layout", [1, 2, 3, 4])
def test_literal_equals_generalised_at_npml32(layout):
    nz, nx = 40, 150
    rng = np.random.default_rng(layout)
    vp = rng.uniform(1800, 3000, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2300, (nz, nx)).astype(np.float32)
    P = O.make_params(nx=nx, nz=nz, npml=32, nt=120, flag_layout=layout, spx0=200.0, offset=40.0,
                      offsetmax=700.0, nshot=1)
    a = O.run_shot(P, 1, vp, rho, literal=True)
    b = O.run_shot(P, 1, vp, rho, literal=False)
    assert np.array_equal(a.seis, b.seis) and np.array_equal(a.g1, b.g1) and np.array_equal(a.illum, b.illum)


def test_literal_form_rejects_other_npml():
    vp, rho = _homog(30, 60)
    P = O.make_params(nx=60, nz=30, npml=40, nt=20)
    with pytest.raises(RuntimeError):
        O.run_shot(P, 1, vp, rho, literal=True)


def _recon_error(f, exact, dt=1e-3, t_snap=0.4):
    nz, nx = 70, 120
    rng = np.random.default_rng(3)
    vp = rng.uniform(1900, 2600, (nz, nx)).astype(np.float32)
    rho = np.full_like(vp, 2000.0)
    nt = int(round(2.5 / dt))
    P = O.make_params(nx=nx, nz=nz, npml=30, nt=nt, dt=dt, fdom=f, spx0=600.0, spz0=200.0, flag_layout=2)
    r = O.run_shot(P, 1, vp, rho, snap_it=int(round(t_snap / dt)), f64=True, exact_reverse=exact)
    n = P.npml
    return rel_l2(r.snap_rec[n:-n, n:-n], r.snap_fwd[n:-n, n:-n])


def test_reconstruction_inverts_forward_pass():
    """With the source removed at it+1 and the strips left as saved, the
    boundary-saved reverse pass inverts the forward pass (the wavefield has
    left the grid by nt, so the zero start of the reverse velocities costs
    nothing): any strip index or stagger mistake would show as O(1)."""
    assert _recon_error(10.0, exact=True) < 1e-3
    assert _recon_error(10.0, exact=True, t_snap=0.08) < 1e-3


def test_reconstruction_residual_is_the_source_lag():
    """Reference semantics (rtm_model.cu:366-437): the source is removed one
    step late, so while the wavelet is on the reconstruction carries an error
    of relative size ~ omega*dt, halved with dt; after the wavelet it is gone."""
    e1 = _recon_error(10.0, exact=False, dt=1e-3, t_snap=0.08)
    e2 = _recon_error(10.0, exact=False, dt=5e-4, t_snap=0.08)
    assert 1e-3 < e1 < 0.5 and 1.8 < e1 / e2 < 2.2, (e1, e2)
    assert _recon_error(10.0, exact=False, t_snap=0.4) < 1e-3


def test_fp32_restatement_vs_fp64():
    nz, nx = 60, 120
    rng = np.random.default_rng(7)
    vp = rng.uniform(1800, 3000, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2300, (nz, nx)).astype(np.float32)
    P = O.make_params(nx=nx, nz=nz, npml=24, nt=400, flag_layout=2, spx0=600.0)
    a = O.run_shot(P, 1, vp, rho)
    b = O.run_shot(P, 1, vp, rho, f64=True)
    assert rel_l2(a.seis, b.seis) < 1e-5
    assert rel_l2(a.g1, b.g1) < 1e-4


def test_kernel_order_form_within_tolerance():
    """the kernel-order FP32 form (1/rho, rho*v^2, damping ratios rounded once)
    stays within the parity tolerances of the reference-order form, and no
    closer to it than rounding allows: it is a different operation order"""
    nz, nx = 60, 120
    rng = np.random.default_rng(11)
    vp = rng.uniform(1800, 3000, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2300, (nz, nx)).astype(np.float32)
    P = O.make_params(nx=nx, nz=nz, npml=24, nt=500, flag_layout=2, spx0=600.0)
    a = O.run_shot(P, 1, vp, rho)
    k = O.run_shot(P, 1, vp, rho, korder=True)
    d = O.run_shot(P, 1, vp, rho, f64=True)
    assert rel_l2(k.seis, a.seis) < 1e-5 and rel_l2(k.g1, a.g1) < 1e-4 and rel_l2(k.illum, a.illum) < 1e-4
    assert 0 < rel_l2(k.seis, a.seis)
    # against FP64 the two FP32 orders are equally close
    ek, ea = rel_l2(k.seis, d.seis), rel_l2(a.seis, d.seis)
    assert ek < 3 * ea and ea < 3 * ek, (ek, ea)


def test_kernel_order_damping_reduces_off_the_layer():
    """one step from random fields: the kernel-order form agrees with the
    reference order to rounding in every field, the PML split pair included,
    and is not identical to it (a different operation order)"""
    nz, nx, npml = 40, 60, 16
    rng = np.random.default_rng(12)
    vp = rng.uniform(1800, 3000, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2300, (nz, nx)).astype(np.float32)
    P = O.make_params(nx=nx, nz=nz, npml=npml, nt=10, flag_layout=2, spx0=300.0)
    g = O.geometry(P, 1)
    F = O.alloc_fields(g, npml)
    for f in F.values():
        f[...] = rng.standard_normal(f.shape).astype(np.float32)
    A = {k: v.copy() for k, v in F.items()}
    O.fd_steps(P, 1, vp, rho, F, 1, korder=True)
    O.fd_steps(P, 1, vp, rho, A, 1)
    for k in F:
        assert rel_l2(F[k], A[k]) < 1e-6, k
    assert not all(np.array_equal(F[k], A[k]) for k in F)


def test_ricker_closed_form():
    nt, f, dt = 300, 20.0, 1e-3
    w = O.ricker(nt, f, dt).astype(np.float64)
    t = np.arange(nt) * dt
    tau = np.pi * f * (t - 1 / f)
    assert np.max(np.abs(w - (1 - 2 * tau ** 2) * np.exp(-tau ** 2))) < 1e-6
    assert np.argmax(w) == round(1 / f / dt)


def test_taper_closed_form():
    nxl, nwin = 301, 50
    t = O.taper_row(nxl, nwin).astype(np.float64)
    ix = np.arange(nxl)
    want = np.ones(nxl)
    want[:nwin] = np.exp(-0.0036 * (nwin - ix[:nwin]) ** 2)
    want[nxl - nwin:] = np.exp(-0.0036 * (ix[nxl - nwin:] - nxl + nwin + 1) ** 2)
    assert np.max(np.abs(t - want)) < 1e-6
    assert np.allclose(t, t[::-1], atol=1e-7)


def test_mute_cut_matches_traveltime():
    nz, nx, v = 60, 200, 2000.0
    vp, rho = _homog(nz, nx, v)
    P = O.make_params(nx=nx, nz=nz, npml=20, nt=3000, dt=1e-3, fdom=20.0, direct=1.5, direct0=0.05,
                      flag_layout=2, spx0=1000.0, spz0=20.0)
    g = O.geometry(P, 1)
    cut = O.mute_cuts(P, g, vp)
    ir = np.arange(g.nrec)
    dist = np.hypot((ir + P.npml - g.nsx) * P.dx, (P.npml - g.nsz) * P.dz)
    nw = int(P.direct / (P.fdom * P.dt))
    t_direct = np.floor(dist * 0.964 / v / P.dt) + nw
    want = np.clip(np.maximum(t_direct, min(int(P.direct0 / P.dt), P.nt)), None, P.nt)
    assert np.max(np.abs(cut - want)) <= 1


def test_laplacian_textbook():
    rng = np.random.default_rng(1)
    img = rng.standard_normal((20, 30)).astype(np.float32)
    out = O.laplacian(img, 7, 5)
    a = img.astype(np.float64)
    want = np.zeros_like(a)
    want[1:-1, 1:-1] = (a[1:-1, 2:] + a[1:-1, :-2] - 2 * a[1:-1, 1:-1]) / 49.0 + \
                       (a[2:, 1:-1] + a[:-2, 1:-1] - 2 * a[1:-1, 1:-1]) / 25.0
    assert np.max(np.abs(out - want)) < 1e-5 * np.max(np.abs(want))
    assert np.all(out[0] == 0) and np.all(out[-1] == 0) and np.all(out[:, 0] == 0) and np.all(out[:, -1] == 0)


def _ic_fields(nzp=24, nxp=26, **val):
    """constant fields of the reference shapes"""
    full = lambda v: np.full((nzp, nxp), v, np.float32)
    return dict(ps=full(val.get("ps", 0.0)), vxs=np.full((nzp, nxp - 1), val.get("vxs", 0.0), np.float32),
                vzs=np.full((nzp - 1, nxp), val.get("vzs", 0.0), np.float32), p=full(val.get("p", 0.0)),
                vxr=np.full((nzp, nxp - 1), val.get("vxr", 0.0), np.float32),
                vzr=np.full((nzp - 1, nxp), val.get("vzr", 0.0), np.float32),
                ps_pre=full(val.get("ps_pre", 0.0)), p_pre=full(val.get("p_pre", 0.0)), vv=full(val.get("vv", 4.0e6)))


def test_imaging_conditions_decomposition_signs():
    """kernel.cu:1417-1450: with the source field moving one way in z and x
    (-ps*vz < 0, -ps*vx < 0) and the receiver field the other way in z only, the
    up-down product alone collects ps*p"""
    f = _ic_fields(ps=2.0, vxs=1.0, vzs=1.0, p=3.0, vxr=1.0, vzr=-1.0)
    r = O.ic_cells(**f, npml=6, dx=10.0, dz=10.0, dt=1e-3)
    assert np.all(r["ud"] == 6.0)
    assert not r["du"].any() and not r["lr"].any() and not r["rl"].any()
    f = _ic_fields(ps=2.0, vxs=-1.0, vzs=-1.0, p=3.0, vxr=1.0, vzr=1.0)   # every direction reversed
    r = O.ic_cells(**f, npml=6, dx=10.0, dz=10.0, dt=1e-3)
    assert np.all(r["du"] == 6.0) and np.all(r["rl"] == 6.0) and not r["ud"].any() and not r["lr"].any()


@pytest.mark.parametrize("nxp", [26, 27])
def test_imaging_conditions_vx_read_at_the_p_index(nxp):
    """kernel.cu:1431,1468 index the vx arrays (rows of nxp - 1) with the
    pressure's flat index iz*nxp + ix.  With vxs alternating in sign along its
    flat storage and vxr = -vxs, the left-right product collects ps*p where the
    element read is positive and the right-left product where it is negative,
    so every cell shows which element it read: the parity of iz*nxp + ix, not
    that of the element [iz][ix] (iz*(nxp - 1) + ix), which differs on every
    other row"""
    nzp, npml = 24, 6
    f = _ic_fields(nzp, nxp, ps=2.0, p=3.0)
    alt = np.where(np.arange(nzp * (nxp - 1)) % 2 == 0, 1.0, -1.0).astype(np.float32).reshape(nzp, nxp - 1)
    f["vxs"], f["vxr"] = alt, -alt
    r = O.ic_cells(**f, npml=npml, dx=10.0, dz=10.0, dt=1e-3)
    iz, ix = np.meshgrid(np.arange(npml, nzp - npml), np.arange(npml, nxp - npml), indexing="ij")
    even = (iz * nxp + ix) % 2 == 0
    assert np.array_equal(r["lr"], np.where(even, 6.0, 0.0).astype(np.float32))
    assert np.array_equal(r["rl"], np.where(even, 0.0, 6.0).astype(np.float32))
    assert not r["ud"].any() and not r["du"].any()
    assert not np.array_equal(even, (iz * (nxp - 1) + ix) % 2 == 0)


def test_imaging_conditions_poynting_angle_bin():
    """kernel.cu:1452-1488 with ps = p = 1: source vector (vxs, vzr), receiver
    vector (vxr, vzr) at cos = 0.3, so theta = pi/2 - acos(0.3)/2 = 0.9377 rad
    falls in bin 53 of width pi/180; g31 += cos(theta), g32 += 16 pi v / sin(theta)"""
    c = (-2.0 + np.sqrt(4.0 - 4 * 0.82 * 0.82)) / (2 * 0.82)      # (c + 1)/sqrt(2 (c^2 + 1)) = 0.3
    f = _ic_fields(ps=1.0, vxs=1.0, p=1.0, vxr=c, vzr=1.0, vv=4.0e6)
    r = O.ic_cells(**f, npml=6, dx=10.0, dz=10.0, dt=1e-3)
    th = np.pi / 2 - 0.5 * np.arccos(0.3)
    assert int(th / (np.pi / 180)) == 53
    assert np.allclose(r["g31"][53], np.cos(th), rtol=1e-5, atol=0)
    assert np.allclose(r["g32"][53], 16 * np.pi * 2000.0 / np.sin(th), rtol=1e-5, atol=0)
    others = np.delete(np.arange(61), 53)
    assert not r["g31"][others].any() and not r["g32"][others].any()
    # parallel vectors: theta = pi/2, outside (0, pi/3]: no contribution
    r = O.ic_cells(**_ic_fields(ps=1.0, vxs=1.0, p=1.0, vxr=1.0, vzr=1.0), npml=6, dx=10.0, dz=10.0, dt=1e-3)
    assert not r["g31"].any() and not r["g32"].any()


def test_imaging_conditions_energy_norm_linear_fields():
    """kernel.cu:1489-1529: for ps = a x + b z and p = c x + d z the central
    differences are exact, so g4 = -(dps)(dp)/(dt^2 v^2) + a c + b d"""
    nzp, nxp, dx, dz, dt = 24, 26, 10.0, 7.5, 1e-3
    z, x = np.meshgrid(np.arange(nzp) * dz, np.arange(nxp) * dx, indexing="ij")
    a, b, c, d = 0.25, -0.5, 0.75, 0.125
    f = _ic_fields(nzp, nxp, vv=2.25e6)
    f["ps"] = (a * x + b * z).astype(np.float32)
    f["p"] = (c * x + d * z).astype(np.float32)
    f["ps_pre"] = f["ps"] - np.float32(0.5)
    f["p_pre"] = f["p"] + np.float32(0.25)
    r = O.ic_cells(**f, npml=6, dx=dx, dz=dz, dt=dt)
    want = -(0.5 * -0.25) / (dt * dt * 2.25e6) + a * c + b * d
    assert np.allclose(r["g4"], want, rtol=2e-5, atol=0)
