"""Field-data host steps on the CPU: the oracle's restatement of the gather
scan, the gather read and the seismogram interpolation (Interp_function.cpp)
pinned by construction and closed forms, and librtm_hip's host functions
(rtmh_scan_gather, rtmh_read_gather, rtmh_interp_seis, rtmh_gather_geometry)
against it.  No GPU compute here."""
import math

import numpy as np
import pytest
import scipy.signal

import gubingluo_amd as G
from oracle import field as F
from oracle import oracle as O


def gather_bytes(shots, nt0, dt_us=2000, big=False, sgy=False, seed=0, dstep=25):
    """SU/SEG-Y gather: shots = [(sx, offset, ntrace)], traces in file order,
    trace i of a shot at offset + i*dstep; returns (bytes, data[list of nt0 x ntrace])"""
    rng = np.random.default_rng(seed)
    e = ">" if big else "<"
    hdr = np.dtype([("tracl", e + "i4"), ("pad0", "V32"), ("offset", e + "i4"), ("pad1", "V32"),
                    ("sx", e + "i4"), ("pad2", "V38"), ("ns", e + "i2"), ("dt", e + "i2"), ("pad3", "V122")])
    assert hdr.itemsize == 240
    out = [b"\0" * 3600] if sgy else []
    data = []
    k = 1
    for sx, off, ntr in shots:
        d = rng.standard_normal((nt0, ntr)).astype(np.float32)
        data.append(d)
        for i in range(ntr):
            h = np.zeros(1, hdr)
            h["tracl"], h["offset"], h["sx"], h["ns"], h["dt"] = k, off + dstep * i, sx, nt0, dt_us
            out.append(h.tobytes())
            out.append(d[:, i].astype(e + "f4").tobytes())
            k += 1
    return b"".join(out), data


SHOTS = [(100, 0, 5), (250, 20, 7), (400, 40, 4)]


@pytest.mark.parametrize("big,sgy", [(False, False), (True, False), (False, True), (True, True)])
def test_scan_gather_by_construction(tmp_path, big, sgy):
    raw, _ = gather_bytes(SHOTS, 30, big=big, sgy=sgy)
    tab, e = F.scan_gather(raw, int(sgy))
    assert e == int(big)
    assert tab[0].tolist() == [100, 250, 400] and tab[1].tolist() == [5, 7, 4]
    assert tab[2].tolist() == [1, 6, 13] and tab[3].tolist() == [0, 20, 40]
    p = tmp_path / "g.su"
    p.write_bytes(raw)
    lt, le = G.scan_gather(str(p), int(sgy))
    assert le == e and np.array_equal(lt, tab)


def test_scan_gather_table_overflow_is_an_error(tmp_path):
    raw, _ = gather_bytes(SHOTS, 10)
    p = tmp_path / "g.su"
    p.write_bytes(raw)
    with pytest.raises(RuntimeError):
        G.scan_gather(str(p), 0, max_shots=2)


@pytest.mark.parametrize("big", [False, True])
def test_read_gather_returns_the_traces(tmp_path, big):
    raw, data = gather_bytes(SHOTS, 30, big=big, sgy=True)
    p = tmp_path / "g.sgy"
    p.write_bytes(raw)
    o = F.read_gather(raw, 6, 7, 30, 9, sgyflag=1, endianflag=int(big))
    assert np.array_equal(o[:, :7], data[1]) and not o[:, 7:].any()     # columns past ntrace stay zero
    lib = G.read_gather(str(p), 6, 7, 30, 9, sgyflag=1, endianflag=int(big))
    assert np.array_equal(lib, o)
    assert np.array_equal(G.read_gather(str(p), 6, 7, 30, 5, 1, int(big)), data[1][:, :5])   # extra traces dropped


def test_kaiser_window_is_scipys():
    """modZeroBessel's 19-term series and the Kaiser formula against scipy's window"""
    ref = scipy.signal.windows.kaiser(21, 2.0)
    assert np.allclose(F.kaiser21(), ref, rtol=0, atol=1e-7)
    assert F.kaiser21()[10] == 1.0


def test_sinc_halving_dt_keeps_the_input_samples():
    """dt0 = 2 dt: output sample 2m sits on input sample m with the window's
    centre weight 1 and every other sinc at an integer (zero): out[2m] = in[m];
    the samples between interpolate a slow sinusoid closely"""
    nt0, nx = 120, 3
    t0 = np.arange(nt0)
    sig = np.stack([np.sin(2 * np.pi * t0 / 40 + ph) for ph in (0.0, 0.7, 1.9)], 1).astype(np.float32)
    out = F.sinc_interp(sig, nx, 2 * nt0 - 40, 0.001, 0.002)
    for m in range(12, nt0 - 25):
        # float32 sample times leave the other sincs ~1e-7 off their zeros
        assert np.allclose(out[2 * m], sig[m], rtol=0, atol=1e-5)
    for m in range(12, nt0 - 25):
        exact = np.sin(2 * np.pi * (m + 0.5) / 40 + np.array([0.0, 0.7, 1.9]))
        assert np.max(np.abs(out[2 * m + 1] - exact)) < 5e-2


def test_sinc_same_rate_weights_by_the_shifted_window():
    """dt = dt0: the window is centred one sample late ((it+1)*dt/dt0), so the
    sample at t = it*dt takes Kaiser weight w[9], not the centre's 1"""
    rng = np.random.default_rng(3)
    sig = rng.standard_normal((80, 2)).astype(np.float32)
    out = F.sinc_interp(sig, 2, 80, 0.002, 0.002)
    w9 = scipy.special.i0(2.0 * math.sqrt(1 - 0.01)) / scipy.special.i0(2.0)
    for it in range(12, 60):
        assert np.allclose(out[it], w9 * sig[it], rtol=0, atol=2e-5)   # w[10] = 1 would miss by ~7e-3 |x|


def test_fk_single_factor_closed_form():
    """lx = 1, a cosine at bin k constant over traces: the spectrum passes with
    H = 1/(1+1e-5) and the mirrored half lands one bin off (rows (nt0+1)/2 down),
    so the output is A (cos(2 pi k t/n) + cos(2 pi (k-1) t/n)), A = nx0 n/2/(1+1e-5)"""
    nt0, nx0, k = 64, 6, 5
    t = np.arange(nt0)
    sig = np.repeat(np.cos(2 * np.pi * k * t / nt0)[:, None], nx0, 1).astype(np.float32)
    out = F.fk_interp(sig, nx0, 1)
    A = nx0 * nt0 / 2 / (1 + 1e-5)
    exp = A * (np.cos(2 * np.pi * k * t / nt0) + np.cos(2 * np.pi * (k - 1) * t / nt0))
    assert np.allclose(out, exp[:, None], rtol=0, atol=1e-6 * A)


@pytest.mark.parametrize("lx,nx0", [(2, 8), (3, 9), (4, 8), (2, 7), (3, 8), (3, 10)])
def test_fk_refinement_closed_form(lx, nx0):
    """lx >= 2, an event constant over traces (a cosine at bin k): its f-k
    spectrum lives at kx = 0 only, C repeats it at kx = m nx0, and the spectrum
    of the zero-padded data there is the box sum D_m = sum_{x < nx0}
    exp(-2 pi i m x / lx) times one time factor, which cancels in H = S / Z.
    Z averages the lx aliases: sum_m D_m = lx ceil(nx0 / lx).  So fine trace x
    gets sum_m H_m exp(2 pi i m x / lx) = lx N(x mod lx) / ceil(nx0 / lx) times
    the lx = 1 output (pinned above), N(r) = #{x' < nx0: x' = r mod lx}: lx on
    every trace when lx divides nx0, and alternating levels when it does not
    (H stays below lx^2, so the clip is not reached; the 1e-5 regularisation
    is the tolerance)"""
    nt0, k = 64, 5
    t = np.arange(nt0)
    sig = np.repeat(np.cos(2 * np.pi * k * t / nt0)[:, None], nx0, 1).astype(np.float32)
    base = F.fk_interp(sig, nx0, 1)[:, :1].astype(np.float64)
    out = F.fk_interp(sig, lx * nx0, lx).astype(np.float64)
    assert out.shape == (nt0, lx * nx0)
    ceil = -(-nx0 // lx)
    for x in range(lx * nx0):
        n = len(range(x % lx, nx0, lx))
        exp = lx * n / ceil * base[:, 0]
        assert np.max(np.abs(out[:, x] - exp)) <= 1e-3 * np.max(np.abs(exp)), (x, n)


@pytest.mark.parametrize("lx,nx0,x1,x2,eps_clip,eps_free", [(2, 8, 5, 2, 0.18, 0.36), (3, 8, 4, 3, 0.095, 0.13)])
def test_fk_clip_closed_form(lx, nx0, x1, x2, eps_clip, eps_free):
    """The |H| > lx^2 clip (Interp_function.cpp:205-207), on two spikes at the
    same time t0: trace x1 (x1 != 0 mod lx, amplitude 1) and trace x2
    (x2 = 0 mod lx, amplitude eps).  The lx aliases of trace x1 cancel in Z, so
    Z = eps exp(-2 pi i kk x2 / Nx) at every wavenumber kk of Nx = lx nx0, the
    time factor cancels in S / Z, and H(kk) = 1 + exp(-2 pi i kk (x1 - x2) / Nx)
    / eps, with |H| between (1 - eps) / eps and (1 + eps) / eps.
    - eps below 1 / (lx^2 + 1): every |H| > lx^2, H = lx everywhere, and the
      output is lx^2 times the lx = 1 output q of trace x1 at fine trace lx x1,
      eps lx^2 q at lx x2, zero elsewhere;
    - eps above 1 / (lx^2 - 1): no |H| reaches lx^2, H is the inverse
      transform of delta(0) + delta(x1 - x2) / eps, so the replicated spectrum
      (the data spread to fine traces lx m) comes back plus a copy shifted by
      x1 - x2 and scaled by 1 / eps: lx q at lx x1, eps lx q at lx x2, lx q / eps
      at lx x1 + x1 - x2, lx q at lx x2 + x1 - x2, zero elsewhere.
    The two eps bracket the threshold: (1 + 0.13) / 0.13 = 8.7 < 9 < 9.5 =
    (1 - 0.095) / 0.095 for lx = 3, so a threshold of lx or 2 lx, a clip to lx^2
    or to a complex value, or no clip fails; the 1e-5 regularisation of Z is the
    tolerance.  Both the oracle and the library."""
    nt0, t0 = 64, 20
    nx = lx * nx0

    def spikes(eps):
        a = np.zeros((nt0, nx0), np.float32)
        a[t0, x1], a[t0, x2] = 1.0, eps
        return a

    for name, fk in (("oracle", lambda a, n, l: F.fk_interp(a, n, l)),
                     ("library", lambda a, n, l: G.interp_seis(a, n, nt0, 0.002, 0.002))):
        for eps, clipped in ((eps_clip, True), (eps_free, False)):
            a = spikes(eps)
            q = np.asarray(F.fk_interp(a, nx0, 1), np.float64)[:, x1]
            out = np.asarray(fk(a, nx, lx), np.float64)
            assert out.shape == (nt0, nx)
            exp = np.zeros((nt0, nx))
            if clipped:
                exp[:, lx * x1] = lx * lx * q
                exp[:, lx * x2] = lx * lx * eps * q
            else:
                exp[:, lx * x1] = lx * q
                exp[:, lx * x2] = lx * eps * q
                exp[:, lx * x1 + x1 - x2] = lx * q / eps
                exp[:, lx * x2 + x1 - x2] = lx * q
            err = np.max(np.abs(out - exp)) / np.max(np.abs(exp))
            assert err <= 1e-4, (name, eps, err)


def test_interp_dispatch_and_copy():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((50, 8)).astype(np.float32)
    assert np.array_equal(F.interp_seis2d(a, 8, 50, 0.002, 0.002), a)        # neither axis refined: copy
    assert np.array_equal(F.interp_seis2d(a, 6, 40, 0.002, 0.002), a[:40, :6])
    assert np.array_equal(F.interp_seis2d(a, 8, 90, 0.001, 0.002), F.sinc_interp(a, 8, 90, 0.001, 0.002))


@pytest.mark.parametrize("nx0,nx,nt0,nt", [(8, 8, 60, 117), (8, 6, 60, 117), (8, 8, 60, 60), (9, 17, 40, 80),
                                           (9, 17, 40, 30), (5, 13, 32, 64)])
def test_library_interp_vs_oracle(nx0, nx, nt0, nt):
    rng = np.random.default_rng(nx0 * 100 + nx)
    a = rng.standard_normal((nt0, nx0)).astype(np.float32)
    dt0 = 0.002
    dt = dt0 * nt0 / nt if nt > nt0 else dt0
    lib = G.interp_seis(a, nx, nt, dt, dt0)
    ref = F.interp_seis2d(a, nx, nt, dt, dt0)
    assert lib.shape == ref.shape == (nt, nx)
    if nx <= nx0:
        assert np.array_equal(lib, ref)                                       # sinc / copy: float ops in one order
    else:
        assert np.max(np.abs(lib - ref)) <= 1e-5 * np.max(np.abs(ref))      # FFTs of two libraries


def test_library_interp_refusals():
    a = np.ones((41, 9), np.float32)
    with pytest.raises(ValueError):
        G.interp_seis(a, 17, 41, 0.002, 0.002)                               # odd nt0 with x-interpolation
    with pytest.raises(ValueError):
        G.interp_seis(np.ones((40, 10), np.float32), 36, 40, 0.002, 0.002)   # lx = 35//9 = 3: 36 > 3*10 traces


def test_gather_geometry_vs_oracle_and_layout3():
    kw = dict(nx=200, nz=40, nt=100, dx=10.0, dz=10.0, dt=1e-3, npml=32, flag_layout=3, mode=1)
    rng = np.random.default_rng(5)
    for _ in range(40):
        sx = float(rng.uniform(0, 900))
        off = float(rng.integers(0, 8)) * 10.0 + float(rng.uniform(0, 9))
        om = off + float(rng.uniform(100, 800))
        P = G.make_params(**kw)
        Po = O.make_params(**dict(kw, field=1, fsx=sx, foffset=off, foffsetmax=om))
        g = G.gather_geometry(P, 1, (sx, off, om))
        go = O.geometry(Po, 1)
        for f in ("nx1", "nx2", "nxl", "nxp", "nsx", "nsz", "nrec", "rcol0", "noffset"):
            assert getattr(g, f) == getattr(go, f), f
    # on grid points the gather's window is layout 3's
    for sx, om in ((300.0, 600.0), (0.0, 990.0), (1200.0, 500.0)):
        P3 = G.make_params(**dict(kw, spx0=sx, dspx=0.0, offset=20.0, offsetmax=om, nx=300))
        g3 = G.geometry(P3, 1)
        gf = G.gather_geometry(G.make_params(**dict(kw, nx=300)), 1, (sx, 20.0, om))
        for f in ("nx1", "nx2", "nxl", "nsx", "nrec", "rcol0", "noffset"):
            assert getattr(g3, f) == getattr(gf, f), f
    with pytest.raises(ValueError):
        G.gather_geometry(G.make_params(**kw), 1, (1800.0, 0.0, 500.0))      # leaves the model
