"""The oracle's restatement of the library's 16-bit boundary store
(orc_store16_block: one scale per step and strip block, binary16 / bfloat16
codes rounded to nearest even, read back as code * scale), pinned against
independent conversions: numpy's float32 -> float16 cast and torch's
float32 -> bfloat16 cast (both IEEE round to nearest even), and against what
the scale rule fixes exactly (the block's largest magnitude becomes +-32768,
an all-zero block keeps scale 1, a denormal maximum takes FLT_MIN, NaN does
not raise the maximum).  A wrong rounding mode, a truncation, a scale off by
a factor or a lost subnormal fails one of these.  Then whole shots with the
16-bit store, and with 16-bit snapshots, stay within the format's rounding of
the exact images."""
import numpy as np
import pytest

from oracle import oracle as O

FLT_MIN = np.float32(1.17549435e-38)


def block_with_unit_scale(rng, n):
    """values whose largest magnitude is exactly 32768, so that scale = 1 and
    the codes are the plain conversions of the values"""
    mags = 2.0 ** rng.uniform(-30, 15, n)                      # subnormal halves up to 32768
    v = (mags * rng.choice([-1.0, 1.0], n)).astype(np.float32)
    v[:7] = [32768.0, -32768.0, 0.0, -0.0, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 2.0 ** -25]
    v[7:12] = [6.1035156e-05, 6.0975552e-05, 5.9604645e-08, 2.9802322e-08, 8.940697e-08]   # 2^-14 and subnormal ties
    return v


def test_binary16_codes_are_the_ieee_conversion():
    rng = np.random.default_rng(1)
    v = block_with_unit_scale(rng, 20000)
    back, codes, sc = O.store16_block(v, 1)
    assert sc == 1.0
    ref = v.astype(np.float16)
    assert np.array_equal(codes, ref.view(np.uint16))
    assert np.array_equal(back, ref.astype(np.float32))
    # the ties to even: 1 + 2^-11 rounds down, 1 + 3*2^-11 up
    assert back[4] == 1.0 and back[5] == np.float32(1.0 + 2.0 ** -9)
    # the subnormal range of binary16 is kept (2^-24 is its smallest step)
    assert back[9] == np.float32(2.0 ** -24) and back[6] == 0.0


def test_bfloat16_codes_are_the_rounded_top_half():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    v = block_with_unit_scale(rng, 20000)
    back, codes, sc = O.store16_block(v, 2)
    assert sc == 1.0
    t = torch.from_numpy(v.copy()).to(torch.bfloat16)
    assert np.array_equal(codes, t.view(torch.int16).numpy().view(np.uint16))
    assert np.array_equal(back, t.to(torch.float32).numpy())


@pytest.mark.parametrize("fmt", [1, 2])
def test_block_scale(fmt):
    rng = np.random.default_rng(3)
    v = rng.normal(0.0, 3.0e-4, 5000).astype(np.float32)
    i = int(np.argmax(np.abs(v)))
    back, codes, sc = O.store16_block(v, fmt)
    assert sc == np.float32(np.abs(v[i]) * np.float32(1.0 / 32768.0))
    # the largest magnitude becomes +-32768 exactly and reads back exactly
    top = 0x7800 if fmt == 1 else 0x4700
    assert codes[i] & 0x7FFF == top and back[i] == v[i]
    # every value reads back within half a code step of the format
    rel = 2.0 ** -11 if fmt == 1 else 2.0 ** -8
    assert np.all(np.abs(back - v) <= rel * np.maximum(np.abs(v), 2.0 ** -14 * sc) + 1e-45)
    # all zero: scale 1, codes 0
    z, zc, zs = O.store16_block(np.zeros(100, np.float32), fmt)
    assert zs == 1.0 and not zc.any() and not z.any()
    # a denormal maximum: the scale floors at FLT_MIN (never zero)
    d, dc, ds = O.store16_block(np.array([1e-40, -3e-41, 0.0], np.float32), fmt)
    assert np.float32(ds) == FLT_MIN and np.all(np.isfinite(d))
    # NaN does not raise the maximum
    nb = np.array([np.nan, 2.0, -1.0], np.float32)
    _, _, ns = O.store16_block(nb, fmt)
    assert ns == np.float32(2.0 / 32768.0)


@pytest.mark.parametrize("fmt,tol", [(1, 2e-4), (2, 2e-3)])
def test_shot_with_16bit_store(fmt, tol):
    """the 16-bit store only perturbs what the reverse pass reads at the strips:
    images within the format's rounding of the exact store's, and different"""
    nz, nx = 60, 200
    rng = np.random.default_rng(4)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[30:] = 2500.0
    vp += rng.uniform(-50, 50, (nz, nx)).astype(np.float32)
    rho = np.full((nz, nx), 2000.0, np.float32)
    P = O.make_params(nshot=1, startshot=1, dshot=1, flag_layout=2, nx=nx, nz=nz, nt=400, dx=10.0, dz=10.0,
                      dt=1e-3, npml=20, fdom=20.0, spx0=1000.0, spz0=20.0, offsetmax=1000.0)
    a = O.run_shot(P, 1, vp, rho, korder=True)
    b = O.run_shot(P, 1, vp, rho, korder=True, store_fmt=fmt)
    assert np.array_equal(a.seis, b.seis)                  # the forward pass does not read the store
    for x, y in ((a.g1, b.g1), (a.illum, b.illum)):
        e = np.linalg.norm((x - y).astype(np.float64)) / np.linalg.norm(x.astype(np.float64))
        assert 0 < e <= tol, e
    with pytest.raises(RuntimeError):
        O.run_shot(P, 1, vp, rho, f64=True, store_fmt=fmt)
    # 16-bit snapshots (one scale per snapshot and field): the restores move
    # the images by the format's rounding only
    c = O.run_shot(P, 1, vp, rho, korder=True, snap_every=50)
    d = O.run_shot(P, 1, vp, rho, korder=True, snap_every=50, snap_fmt=fmt)
    e = np.linalg.norm((c.g1 - d.g1).astype(np.float64)) / np.linalg.norm(c.g1.astype(np.float64))
    assert 0 < e <= tol, e
    with pytest.raises(RuntimeError):
        O.run_shot(P, 1, vp, rho, f64=True, snap_every=50, snap_fmt=fmt)
