"""Write the golden fixtures of tests/golden/ from the reference's own host code.

The reference ships no tests or fixtures.  Its host helpers compile from
their own sources (oracle/Makefile `ref` target -> oracle/_ref/libref_host.so,
built only where /root/reference exists); this script calls them through the
extern "C" wrappers of oracle/ref_harness.cpp on small seeded inputs and
stores inputs and outputs as data:

    coef_table.npz        Diff_coeffVelStrLiuyang, orders 2..20 (FD_function.cpp:73-165)
    window_pad.npz        extractvel1/extractrho1 + extendvel1 (Velocity_function.cpp:2011-2073)
    interpvel.npz         interpvel1 (Velocity_function.cpp:1946-1977)
    laplacian.npz         laplacefd1d with int spacings (image_function.cpp:44-79)
    scaleimage.npz        scaleimage1d (image_function.cpp:11-20)
    seis_su_le.bin        Outseisfile1d, SU little endian (IO_function.cpp:131-182)
    seis_sgy_be.bin       Outseisfile1d, SEG-Y (3600-byte volume header) big endian
    seis_inputs.npz       the seismogram and key header written to the two files above
    image_su.bin          Output1d (IO_function.cpp:366-398)
    image_inputs.npz      the image written above
    smoothing.npz         pmlvelsmooth1d / velsmooth1d (Velocity_function.cpp:9-103, 199-313)

Run:  make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_host.so")


def fp(a):
    return a.ctypes.data_as(C.c_void_p)


def main():
    if not os.path.exists(LIB):
        sys.exit(f"missing {LIB}: build it with `make -C oracle ref` where the reference sources exist")
    L = C.CDLL(LIB)
    L.ref_maxval.restype = C.c_float
    L.ref_sum1abs.restype = C.c_float
    rng = np.random.default_rng(20261202)

    # coefficient table
    tab = np.zeros((10, 10), np.float32)
    for r in range(1, 11):
        row = np.zeros(10, np.float32)
        L.ref_diff_coef(fp(row), C.c_int(2 * r))
        tab[r - 1, :r] = row[:r]
    s20 = L.ref_sum1abs(fp(np.ascontiguousarray(tab[9])), C.c_int(10))
    np.savez(os.path.join(HERE, "coef_table.npz"), table=tab, sum1abs_order20=np.float32(s20))

    # window extraction + PML padding
    nz, nx, npml = 12, 30, 6
    vp = rng.uniform(1500, 4500, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1000, 2600, (nz, nx)).astype(np.float32)
    out = {}
    for nx1, nx2 in [(0, nx - 1), (5, 20), (17, 29)]:
        nxl = nx2 - nx1 + 1
        for name, src, sq in (("vv", vp, 1), ("rho", rho, 0)):
            o = np.zeros((nz + 2 * npml, nxl + 2 * npml), np.float32)
            L.ref_window_pad(fp(o), fp(src), C.c_int(nx), C.c_int(nz), C.c_int(nx1), C.c_int(nx2), C.c_int(npml),
                             C.c_int(sq))
            out[f"{name}_{nx1}_{nx2}"] = o
    np.savez(os.path.join(HERE, "window_pad.npz"), vp=vp, rho=rho, npml=npml, **out)

    # model interpolation
    nx0, nz0 = 9, 7
    vp0 = rng.uniform(1500, 4500, (nz0, nx0)).astype(np.float32)
    res = {}
    for lx, lz in [(1, 1), (2, 1), (2, 2), (4, 2)]:
        nxo, nzo = (nx0 - 1) * lx + 1, (nz0 - 1) * lz + 1
        o = np.zeros((nzo, nxo), np.float32)
        L.ref_interpvel(fp(vp0), fp(o), C.c_int(nx0), C.c_int(nz0), C.c_int(nxo), C.c_int(nzo), C.c_int(lx),
                        C.c_int(lz))
        res[f"out_{lx}_{lz}"] = o
    np.savez(os.path.join(HERE, "interpvel.npz"), vp0=vp0, **res)

    # Laplacian (int spacings) and image scaling
    img = rng.standard_normal((11, 13)).astype(np.float32)
    lap = {}
    for dx, dz in [(10, 10), (7, 7), (12, 5)]:
        a = img.copy()
        L.ref_laplacian(C.c_int(13), C.c_int(11), C.c_int(dx), C.c_int(dz), fp(a))
        lap[f"lap_{dx}_{dz}"] = a
    np.savez(os.path.join(HERE, "laplacian.npz"), image=img, **lap)
    ill = np.abs(rng.standard_normal(50)).astype(np.float32)
    grad = rng.standard_normal(50).astype(np.float32)
    g2 = grad.copy()
    L.ref_scaleimage(fp(ill), fp(g2), C.c_int(50))
    np.savez(os.path.join(HERE, "scaleimage.npz"), illum=ill, grad=grad, out=g2)

    # SU / SEG-Y seismogram and image files
    nt, ntr = 37, 5
    seis = rng.standard_normal((nt, ntr)).astype(np.float32)
    key = dict(itraceall=41, itrace=1, offset=-200, sx=1500, gx=1300, nt=nt, dt_us=800, dx=12.5, ntrace=ntr)
    with tempfile.TemporaryDirectory() as td:
        for fname, sgy, endian in (("seis_su_le.bin", 0, 0), ("seis_sgy_be.bin", 1, 1)):
            p = os.path.join(td, "s.su")
            L.ref_write_seis(fp(seis), *(C.c_int(key[k]) for k in ("itraceall", "itrace", "offset", "sx", "gx",
                                                                  "nt", "dt_us")), C.c_float(key["dx"]),
                             C.c_int(ntr), C.c_int(0), C.c_int(sgy), C.c_int(endian), p.encode())
            with open(p, "rb") as f, open(os.path.join(HERE, fname), "wb") as g:
                g.write(f.read())
        np.savez(os.path.join(HERE, "seis_inputs.npz"), seis=seis, **{k: np.array(v) for k, v in key.items()})
        image = rng.standard_normal((9, 6)).astype(np.float32)
        p = os.path.join(td, "i.su")
        L.ref_write_image(fp(image), C.c_int(9), C.c_int(6), C.c_float(10.0), p.encode())
        with open(p, "rb") as f, open(os.path.join(HERE, "image_su.bin"), "wb") as g:
            g.write(f.read())
        np.savez(os.path.join(HERE, "image_inputs.npz"), image=image, d=np.float32(10.0))

    # flag-conditioned model smoothing (off in every benchmark configuration)
    nzp, nxp, npml = 24, 28, 6
    base = rng.uniform(1500, 3000, (nzp, nxp)).astype(np.float32)
    a = base.copy()
    L.ref_pmlvelsmooth(fp(a), C.c_int(nxp), C.c_int(nzp), C.c_int(npml))
    b = base.copy()
    L.ref_velsmooth(fp(b), C.c_int(nzp), C.c_int(nxp), C.c_int(3))
    np.savez(os.path.join(HERE, "smoothing.npz"), base=base, npml=npml, pml=a, nsp=3, smooth=b)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
