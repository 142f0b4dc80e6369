"""rtmh_postprocess_noalloc (the image post-processing with no device
allocation: staging in a file-scope device array) against rtmh_postprocess
through the C ABI: bit for bit on two calls in a row, the second with a
smaller image, and a size past its limit refused."""
import ctypes as C

import numpy as np
import pytest

import gubingluo_amd as G

pytestmark = pytest.mark.gpu

MAX_CELLS = 2048 * 1024


def _call(fn, img, ill, dx, dz):
    nz, nx = img.shape
    lap = np.full_like(img, np.nan)
    nrm = np.full_like(img, np.nan)
    rc = fn(C.c_int(0), G._fp(img), G._fp(ill), C.c_int(nx), C.c_int(nz), C.c_float(dx), C.c_float(dz),
            G._fp(lap), G._fp(nrm))
    return rc, lap, nrm


def test_no_alloc_postprocess_equals_allocating():
    L = G.lib()
    rng = np.random.default_rng(17)
    for nz, nx, dx, dz in ((1024, 2048, 7.5, 7.5), (300, 517, 10.0, 12.0)):   # the limit itself, then smaller
        img = rng.standard_normal((nz, nx)).astype(np.float32)
        ill = rng.uniform(0.0, 3.0, (nz, nx)).astype(np.float32)
        print(f"[no-alloc] {nz}x{nx}", flush=True)
        r1, lap1, nrm1 = _call(L.rtmh_postprocess, img, ill, dx, dz)
        r2, lap2, nrm2 = _call(L.rtmh_postprocess_noalloc, img, ill, dx, dz)
        assert r1 == 0 and r2 == 0, (r1, r2, G.last_error())
        assert np.array_equal(lap1, lap2) and np.array_equal(nrm1, nrm2)
        assert np.isfinite(lap2).all() and np.isfinite(nrm2).all()


def test_no_alloc_postprocess_refuses_past_its_limit():
    big = np.zeros((1025, 2048), np.float32)
    assert big.size > MAX_CELLS
    rc, _, _ = _call(G.lib().rtmh_postprocess_noalloc, big, big, 10.0, 10.0)
    assert rc == -3 and "no-allocation limit" in G.last_error()
