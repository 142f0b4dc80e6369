"""Pins of the oracle's shot windows and absorbing-layer profile by what the
physics fixes, not by a second copy of the formulas (runs on the CPU):

  PML profile     a plane pulse entering each of the four absorbing layers
                  loses amplitude as exp(-int_0^s d(x)/v dx) with
                  d = d0 ((s)/L)^2, d0 = 1.5 vmax ln(1/R)/L, R = 1e-10,
                  L = npml h (kernel.cu:174-269): a d0 off by 2, a linear
                  profile or half nodes off by one cell miss by > 10 %
  windows         layout windows as rtm_model.cu:124-134 (split spread,
                  source in the middle, clipped at the model edges),
                  1125-1155 (one-sided right, receivers from the offset),
                  1632-1660 (one-sided left) on hand-worked cases
  record layout   trace k of a record is the receiver at the model column
                  the window says: the direct wave's earliest arrival is at
                  the source's column and a point scatterer's earliest
                  scattered arrival at the scatterer's column (apex of the
                  traveltime curves, located to a fraction of a trace)
  imaging         a flat reflector migrated shot by shot with layouts 1, 3
                  and 4 images at its depth in every imaged column
"""
import numpy as np
import pytest

from oracle import oracle as O

LN_1_OVER_R = float(np.log(np.float32(1.0) / np.float32(1e-10)))


# ------------------------------------------------------------------ PML --
def _plane_pulses(axis, n_int=200, npml=40, h=10.0, v=2000.0, rho=2000.0, dt=1e-3, sigma=30.0):
    """Two plane pulses (uniform along the other axis) leaving the middle of
    the grid towards both absorbing layers of `axis`; returns the pressure
    traces on the middle line at every cell of that axis (nt, n)."""
    nlong = n_int + 80                       # the other axis: long enough that its edges stay quiet
    nx, nz = (n_int, nlong) if axis == "x" else (nlong, n_int)
    vp = np.full((nz, nx), v, np.float32)
    rh = np.full((nz, nx), rho, np.float32)
    P = O.make_params(nx=nx, nz=nz, npml=npml, nt=4000, dt=dt, dx=h, dz=h, flag_layout=2, spx0=h * nx / 2)
    g = O.geometry(P, 1)
    F = O.alloc_fields(g, npml, f64=True)
    nxp, nzp = g.nxp, g.nzp
    n = nxp if axis == "x" else nzp
    mid = (n - 1) / 2.0
    c_lo, c_hi = mid - 20.0, mid + 20.0      # left/top pulse moves towards 0, the other towards n-1
    i = np.arange(n, dtype=np.float64)
    ih = i + 0.5                             # velocity nodes
    f = lambda c, x: np.exp(-((x - c) * h / sigma) ** 2)
    sh = v * dt / 2 / h                      # velocities lead by half a step (v^{-1/2})
    p1d = f(c_lo, i) + f(c_hi, i)
    v1d = (-f(c_lo, ih - sh) + f(c_hi, ih + sh)) / (rho * v)
    other = nzp if axis == "x" else nxp
    lo, hi = npml, other - npml              # rows (columns) where the pulses start: off the other layers
    if axis == "x":
        F["p"][lo:hi, :] = p1d[None, :]
        F["vx"][lo:hi, :] = v1d[None, :nxp - 1]
    else:
        F["p"][:, lo:hi] = p1d[:, None]
        F["vz"][:, lo:hi] = v1d[:nzp - 1, None]
    line = other // 2
    # until the pulse has passed 26 cells into the layer
    nsteps = int(((n - npml - 1 + 26 - c_hi) * h + 4 * sigma) / (v * dt))
    tr = np.zeros((nsteps, n))
    for it in range(nsteps):
        O.fd_steps(P, 1, vp, rh, F, 1, it0=it, f64=True)
        tr[it] = F["p"][line, :] if axis == "x" else F["p"][:, line]
    # the pulses must not have met anything from the long axis' edges
    assert nsteps * v * dt < (line - lo) * h
    return tr, g, P


@pytest.mark.parametrize("axis", ["x", "z"])
def test_pml_plane_wave_attenuation(axis):
    npml = 40
    tr, g, P = _plane_pulses(axis, npml=npml)
    n = tr.shape[1]
    L = float(npml)
    peak = np.abs(tr).max(axis=0)
    for side in ("low", "high"):
        # the profile starts at the last undamped node (depth 0) and deepens
        # towards the grid edge (kernel.cu:174-269)
        origin = npml if side == "low" else n - npml - 1
        step = -1 if side == "low" else 1
        a0 = peak[origin]
        for s in (10, 20, 24):
            want = np.exp(-0.5 * LN_1_OVER_R * (s / L) ** 3)
            got = peak[origin + step * s] / a0
            assert abs(got / want - 1.0) < 0.03, (axis, side, s, got, want)


def test_pml_attenuation_discriminates_profiles():
    """The plane-wave law above tells the quadratic profile from the
    mistakes it is meant to catch, by more than the 3 % tolerance."""
    k = 0.5 * LN_1_OVER_R
    for s in (20 / 40, 24 / 40):
        right = np.exp(-k * s ** 3)
        assert abs(np.exp(-2 * k * s ** 3) / right - 1) > 0.1          # d0 twice as large
        lin = np.exp(-1.5 * LN_1_OVER_R * s ** 2 / 2)                    # linear profile
        assert abs(lin / right - 1) > 0.1
        # half nodes one cell deeper: the velocity half of the damping lags one cell
        lag = np.exp(-0.5 * k * (s ** 3 + (s - 1 / 40) ** 3))
        assert abs(lag / right - 1) > 0.05


# -------------------------------------------------------------- windows --
def _geo(**kw):
    d = dict(nx=400, nz=50, dx=10.0, dz=10.0, npml=30, nt=100, spx0=1000.0, dspx=100.0, offset=0.0,
             offsetmax=2000.0, nshot=5)
    d.update(kw)
    return d


def test_layout1_windows_hand_cases():
    """offsetmax 2000 m at 10 m: NX = 201 receivers, 100 on each side of the
    source; shots every (int)(dspx/dx) columns from (int)(spx0/dx)."""
    P = O.make_params(flag_layout=1, **_geo())
    for ishot, (nx1, nx2) in {1: (0, 200), 2: (10, 210), 3: (20, 220), 5: (40, 240)}.items():
        g = O.geometry(P, ishot)
        src = 100 + 10 * (ishot - 1)
        assert (g.nx1, g.nx2, g.nxl, g.nrec) == (nx1, nx2, nx2 - nx1 + 1, nx2 - nx1 + 1), ishot
        assert g.nsx - P.npml == src - nx1 and g.rcol0 == P.npml
    # even NX (offsetmax 1990 m: 200 receivers asked) keeps nr = (NX-1)/2 = 99 on each side
    P = O.make_params(flag_layout=1, **_geo(offsetmax=1990.0, spx0=2000.0))
    g = O.geometry(P, 1)
    assert (g.nx1, g.nx2, g.nxl) == (101, 299, 199) and g.nsx - P.npml == 99
    # clipped at both model edges
    P = O.make_params(flag_layout=1, **_geo(spx0=300.0, dspx=3000.0))
    g1, g2 = O.geometry(P, 1), O.geometry(P, 2)
    assert (g1.nx1, g1.nx2) == (0, 130) and g1.nsx - P.npml == 30
    assert (g2.nx1, g2.nx2) == (230, 399) and g2.nsx - P.npml == 330 - 230


def test_layout3_layout4_windows_hand_cases():
    """one-sided spreads of NX-1 = 100 columns (offsetmax 1000 m): layout 3
    from the source to the right with receivers from the offset (50 m = 5
    columns) on; layout 4 from 100 columns left of the source to it, its
    receivers from the window's first column, offset columns short."""
    P = O.make_params(flag_layout=3, **_geo(offsetmax=1000.0, offset=50.0, spx0=1500.0))
    g = O.geometry(P, 2)                     # source column 160
    assert (g.nx1, g.nx2, g.nxl, g.noffset) == (160, 260, 101, 5)
    assert g.nsx - P.npml == 0 and g.rcol0 == P.npml + 5 and g.nrec == 96
    P = O.make_params(flag_layout=4, **_geo(offsetmax=1000.0, offset=50.0, spx0=1500.0))
    g = O.geometry(P, 2)
    assert (g.nx1, g.nx2, g.nxl) == (60, 160, 101)
    assert g.nsx - P.npml == 100 and g.rcol0 == P.npml and g.nrec == 96
    # clipped: layout 3 at the right edge, layout 4 at the left edge
    g = O.geometry(O.make_params(flag_layout=3, **_geo(offsetmax=1000.0, spx0=3500.0)), 1)
    assert (g.nx1, g.nx2) == (350, 399)
    g = O.geometry(O.make_params(flag_layout=4, **_geo(offsetmax=1000.0, spx0=400.0)), 1)
    assert (g.nx1, g.nx2) == (0, 40) and g.nsx - P.npml == 40


# ---------------------------------------------------------- record apex --
def _first_arrivals(rec, frac=0.25):
    """per trace, the (linearly interpolated) sample where |trace| first
    reaches frac of its maximum"""
    out = np.zeros(rec.shape[1])
    for k in range(rec.shape[1]):
        a = np.abs(rec[:, k].astype(np.float64))
        thr = frac * a.max()
        j = int(np.argmax(a >= thr))
        out[k] = j - 1 + (thr - a[j - 1]) / (a[j] - a[j - 1]) if j > 0 else 0.0
    return out


def _apex(t):
    """sub-trace position of the minimum of a traveltime curve (parabola
    through the smallest sample and its neighbours)"""
    k = int(np.argmin(t))
    k = min(max(k, 1), len(t) - 2)
    a, b, c = t[k - 1], t[k], t[k + 1]
    den = a - 2 * b + c
    return k + (0.5 * (a - c) / den if den > 0 else 0.0)


@pytest.mark.parametrize("layout,kw", [
    (1, dict(offsetmax=1600.0, spx0=1200.0, dspx=300.0)),
    (2, dict(spx0=1500.0)),
    (3, dict(offsetmax=1600.0, offset=0.0, spx0=700.0, dspx=300.0)),
    (4, dict(offsetmax=1600.0, offset=0.0, spx0=2100.0, dspx=300.0)),
])
def test_record_traces_sit_at_the_window_columns(layout, kw):
    nz, nx, v = 80, 300, 2000.0
    xd, zd = 155, 40                         # scatterer: model column and depth (cells)
    vp = np.full((nz, nx), v, np.float32)
    vp[-2:] = 3000.0                         # both models: the same vmax, hence the same PML
    rho = np.full((nz, nx), 2000.0, np.float32)
    vp_d = vp.copy()
    vp_d[zd - 1:zd + 2, xd - 1:xd + 2] = 3000.0   # a bulk-modulus scatterer radiates alike in all directions
    P = O.make_params(nx=nx, nz=nz, npml=30, nt=1700, dt=5e-4, fdom=30.0, spz0=20.0, flag_layout=layout, nshot=2,
                      **kw)
    for ishot in (1, 2):
        g = O.geometry(P, ishot)
        a = O.run_shot(P, ishot, vp, rho, forward_only=True, f64=True).seis
        b = O.run_shot(P, ishot, vp_d, rho, forward_only=True, f64=True).seis
        col0 = g.nx1 + (g.rcol0 - P.npml)    # model column of trace 0
        src = g.nx1 + g.nsx - P.npml         # model column of the source
        if 0 <= src - col0 < g.nrec:
            # direct wave: earliest at the source's column (interior apex, or
            # the end trace when the source sits on it)
            t = _first_arrivals(a)
            k = src - col0
            if 0 < k < g.nrec - 1:
                assert abs(_apex(t) - k) < 0.25, (layout, ishot, _apex(t), k)
            else:
                assert int(np.argmin(t)) == k, (layout, ishot)
        if 2 <= xd - col0 < g.nrec - 2:
            # the scattered field alone; the 3x3-cell scatterer's waveform
            # varies with angle, which moves the picks of one-sided spreads by
            # up to a third of a trace (a trace off would be 0.7 or more)
            t = _first_arrivals(b - a, 0.5)
            assert abs(_apex(t) - (xd - col0)) < 0.4, (layout, ishot, _apex(t), xd - col0)


# ------------------------------------------------------------- imaging --
@pytest.mark.parametrize("layout,kw", [
    (1, dict(offsetmax=1200.0, spx0=800.0, dspx=400.0)),
    (3, dict(offsetmax=1200.0, offset=0.0, spx0=300.0, dspx=400.0)),
    (4, dict(offsetmax=1200.0, offset=0.0, spx0=1500.0, dspx=400.0)),
])
def test_flat_reflector_images_at_its_depth(layout, kw):
    nz, nx, zr = 90, 200, 55
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[zr:] = 2800.0
    rho = np.full((nz, nx), 2000.0, np.float32)
    P = O.make_params(nx=nx, nz=nz, npml=30, nt=1100, dt=1e-3, fdom=25.0, spz0=20.0, flag_layout=layout,
                      nshot=3, **kw)
    i1, i2, il = O.rtm(P, vp, rho)
    cols = [ix for ix in range(nx) if np.abs(i1[:, ix]).max() > 0]
    assert len(cols) > 0
    # columns no window covers stay exactly zero
    covered = set()
    for s in (1, 2, 3):
        g = O.geometry(P, s)
        covered.update(range(g.nx1, g.nx2 + 1))
    assert set(cols) <= covered
    # the cross-correlation image peaks at the reflector (within a cell) in
    # every column it images strongly (half its largest value near the depth)
    band = np.abs(i1[zr - 20:zr + 20].astype(np.float64))
    bm = band.max(axis=0)
    strong = [ix for ix in range(nx) if bm[ix] > 0.5 * bm.max()]
    assert len(strong) >= 60, len(strong)
    for ix in strong:
        assert abs(int(np.argmax(band[:, ix])) - 20) <= 1, (layout, ix, int(np.argmax(band[:, ix])) - 20)
