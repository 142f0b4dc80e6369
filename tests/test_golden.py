"""Golden vectors written by the reference's own host helpers
(tests/golden/make_golden.py, from oracle/_ref/libref_host.so built from the
unmodified reference sources) against the oracle's restatements and the
product's host layer (librtm_hip.so, no GPU needed for these functions)."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import gubingluo_amd as G
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_coefficient_table():
    d = load("coef_table.npz")
    assert np.array_equal(O.coef_table().reshape(10, 10), d["table"])
    assert np.array_equal(G.coef_table().reshape(10, 10), d["table"])
    assert abs(float(d["sum1abs_order20"]) - 1.4760199) < 1e-6


def test_window_pad():
    d = load("window_pad.npz")
    vp, rho, npml = d["vp"], d["rho"], int(d["npml"])
    for nx1, nx2 in [(0, vp.shape[1] - 1), (5, 20), (17, 29)]:
        assert np.array_equal(O.window_pad(vp, nx1, nx2, npml, True), d[f"vv_{nx1}_{nx2}"])
        assert np.array_equal(O.window_pad(rho, nx1, nx2, npml, False), d[f"rho_{nx1}_{nx2}"])


def _smooth(lib, fn, w, *args):
    out = np.ascontiguousarray(w, np.float32).copy()
    getattr(lib, fn)(G._fp(out), *[C.c_int(int(x)) for x in args])
    return out


def test_model_smoothing():
    """pmlvelsmooth1d / velsmooth1d outputs of the reference itself: the
    oracle's restatement and the library's host functions equal them."""
    d = load("smoothing.npz")
    base, npml, nsp = d["base"], int(d["npml"]), int(d["nsp"])
    nzp, nxp = base.shape
    for lib, pml, gauss in ((O.lib(), "orc_pml_smooth", "orc_gauss_smooth"),
                            (G.lib(), "rtmh_pml_smooth", "rtmh_gauss_smooth")):
        assert np.array_equal(_smooth(lib, pml, base, nxp, nzp, npml), d["pml"]), pml
        assert np.array_equal(_smooth(lib, gauss, base, nzp, nxp, nsp), d["smooth"]), gauss
    # the absorbing layer is relaxed, the interior untouched; the Gaussian
    # filter preserves a constant window
    pml = d["pml"]
    assert np.array_equal(pml[npml:nzp - npml, npml:nxp - npml], base[npml:nzp - npml, npml:nxp - npml])
    const = np.full((13, 17), 2500.0, np.float32)
    assert np.allclose(_smooth(G.lib(), "rtmh_gauss_smooth", const, 13, 17, 4), const, rtol=1e-6, atol=0)


def test_interp_model():
    d = load("interpvel.npz")
    vp0 = d["vp0"]
    nz0, nx0 = vp0.shape
    for lx, lz in [(1, 1), (2, 1), (2, 2), (4, 2)]:
        want = d[f"out_{lx}_{lz}"]
        out = np.zeros_like(want)
        G.lib().rtmh_interp_model(G._fp(vp0), G._fp(out), C.c_int(nx0), C.c_int(nz0), C.c_int(want.shape[1]),
                                  C.c_int(want.shape[0]), C.c_int(lx), C.c_int(lz))
        assert np.array_equal(out, want), (lx, lz)


def test_laplacian():
    d = load("laplacian.npz")
    for dx, dz in [(10, 10), (7, 7), (12, 5)]:
        assert np.array_equal(O.laplacian(d["image"], dx, dz), d[f"lap_{dx}_{dz}"])


def test_scaleimage():
    d = load("scaleimage.npz")
    assert np.array_equal(O.scaleimage(d["illum"], d["grad"]), d["out"])


@pytest.mark.parametrize("fname,sgy,endian", [("seis_su_le.bin", 0, 0), ("seis_sgy_be.bin", 1, 1)])
def test_seismogram_file_bytes(fname, sgy, endian):
    """Outseisfile1d bytes: the product writer (host_io.cpp) reproduces the
    reference's file exactly, and the product reader reads it back."""

    class SuKey(C.Structure):
        _fields_ = [("iTraceAll", C.c_int32), ("iTrace", C.c_int32), ("Offset", C.c_int32), ("sx", C.c_int32),
                    ("gx", C.c_int32), ("Nt", C.c_int16), ("Dt", C.c_int16), ("Dx", C.c_float),
                    ("Ntrace", C.c_int32)]

    d = load("seis_inputs.npz")
    seis = np.ascontiguousarray(d["seis"])
    k = SuKey(int(d["itraceall"]), int(d["itrace"]), int(d["offset"]), int(d["sx"]), int(d["gx"]), int(d["nt"]),
              int(d["dt_us"]), float(d["dx"]), int(d["ntrace"]))
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "x.su")
        rc = G.lib().rtmio_write_seis(G._fp(seis), k, C.c_int(0), C.c_int(sgy), C.c_int(endian), p.encode())
        assert rc == 0
        got = open(p, "rb").read()
        assert got == open(os.path.join(GOLD, fname), "rb").read()
        back = np.zeros_like(seis)
        k2 = SuKey()
        rc = G.lib().rtmio_read_seis(G._fp(back), C.byref(k2), C.c_int(0), C.c_int(sgy), p.encode(),
                                     C.c_int(seis.size))
        assert rc == 0 and k2.Nt == seis.shape[0] and k2.Ntrace == seis.shape[1]
        assert np.array_equal(back, seis)


def test_image_file_bytes():
    d = load("image_inputs.npz")
    img = np.ascontiguousarray(d["image"])
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "i.su")
        G.lib().rtmio_write_image(G._fp(img), C.c_int(img.shape[0]), C.c_int(img.shape[1]), C.c_float(float(d["d"])),
                                  p.encode())
        assert open(p, "rb").read() == open(os.path.join(GOLD, "image_su.bin"), "rb").read()


REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref_host.so")


@pytest.mark.skipif(not os.path.exists(REF), reason="reference host helpers not built here")
def test_live_cross_check_with_reference_helpers():
    """Where the reference's helpers are built, random cases beyond the
    committed fixtures."""
    L = C.CDLL(REF)
    rng = np.random.default_rng(99)
    for _ in range(5):
        nz, nx = rng.integers(5, 20), rng.integers(8, 40)
        npml = int(rng.integers(10, 16))
        m = rng.uniform(1000, 5000, (nz, nx)).astype(np.float32)
        nx1 = int(rng.integers(0, nx // 2))
        nx2 = int(rng.integers(nx1, nx))
        want = np.zeros((nz + 2 * npml, nx2 - nx1 + 1 + 2 * npml), np.float32)
        L.ref_window_pad(G._fp(want), G._fp(m), C.c_int(nx), C.c_int(nz), C.c_int(nx1), C.c_int(nx2), C.c_int(npml),
                         C.c_int(1))
        assert np.array_equal(O.window_pad(m, nx1, nx2, npml, True), want)
        img = rng.standard_normal((nz, nx)).astype(np.float32)
        a = img.copy()
        L.ref_laplacian(C.c_int(nx), C.c_int(nz), C.c_int(9), C.c_int(11), G._fp(a))
        assert np.array_equal(O.laplacian(img, 9, 11), a)
        # model smoothing on a padded window of random size, layer and radius
        wz, wx, wp, ns = int(rng.integers(22, 40)), int(rng.integers(22, 50)), int(rng.integers(2, 10)), int(rng.integers(1, 6))
        win = rng.uniform(1500, 4500, (wz, wx)).astype(np.float32)
        ref_pml = _smooth(L, "ref_pmlvelsmooth", win, wx, wz, wp)
        ref_gau = _smooth(L, "ref_velsmooth", win, wz, wx, ns)
        for lib, pml, gauss in ((O.lib(), "orc_pml_smooth", "orc_gauss_smooth"),
                                (G.lib(), "rtmh_pml_smooth", "rtmh_gauss_smooth")):
            assert np.array_equal(_smooth(lib, pml, win, wx, wz, wp), ref_pml), (pml, wz, wx, wp)
            assert np.array_equal(_smooth(lib, gauss, win, wz, wx, ns), ref_gau), (gauss, wz, wx, ns)
