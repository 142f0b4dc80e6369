"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker or the CPU baseline.  The product
package (gubingluo_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class Params(C.Structure):
    """Scalar arguments of rtm_model() (rtm_model.cu:12-17), same order."""
    _fields_ = [
        ("nshot", C.c_int), ("startshot", C.c_int), ("dshot", C.c_int), ("mode", C.c_int),
        ("medium_flag", C.c_int), ("flag_layout", C.c_int), ("sgyflag", C.c_int),
        ("endianflag", C.c_int), ("ismth", C.c_int),
        ("nx", C.c_int), ("nz", C.c_int), ("nt", C.c_int),
        ("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
        ("npml", C.c_int),
        ("direct", C.c_float), ("direct0", C.c_float), ("tsnap", C.c_float), ("fdom", C.c_float),
        ("spx0", C.c_float), ("spz0", C.c_float), ("dspx", C.c_float), ("offset", C.c_float),
        ("offsetmax", C.c_float), ("amp", C.c_float), ("alp", C.c_float),
        # a recorded gather's geometry (rtm_real.cu:121-136) when field != 0
        ("field", C.c_int), ("fsx", C.c_float), ("foffset", C.c_float), ("foffsetmax", C.c_float),
    ]


class Geom(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "ishot", "nx1", "nx2", "nxl", "nxp", "nzp", "nsx", "nsz", "nrec", "rcol0", "noffset",
        "hk_itraceall", "hk_offset", "hk_sx", "hk_gx", "hk_ntrace", "hk_nt", "hk_dt")] + [("hk_dx", C.c_float)]


class ShotOpts(C.Structure):
    _fields_ = [("literal", C.c_int), ("forward_only", C.c_int), ("nt_limit", C.c_int),
                ("snap_it", C.c_int), ("vvpad", C.c_void_p), ("rhopad", C.c_void_p), ("exact_reverse", C.c_int),
                ("ic_out", C.c_void_p), ("snap_every", C.c_int), ("store_fmt", C.c_int),
                ("snap_fmt", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_params(**kw) -> Params:
    d = dict(nshot=1, startshot=1, dshot=1, mode=0, medium_flag=0, flag_layout=2, sgyflag=0,
             endianflag=0, ismth=0, nx=100, nz=60, nt=200, dx=10.0, dz=10.0, dt=1e-3, npml=32,
             direct=1.5, direct0=0.05, tsnap=0.1, fdom=20.0, spx0=500.0, spz0=20.0, dspx=50.0,
             offset=0.0, offsetmax=1000.0, amp=1.0, alp=0.05)
    d.update(kw)
    P = Params()
    for k, v in d.items():
        setattr(P, k, v)
    return P


def params_dict(P: Params) -> dict:
    return {n: getattr(P, n) for n, _ in P._fields_}


def geometry(P: Params, ishot: int) -> Geom:
    g = Geom()
    rc = lib().orc_shot_geometry(C.byref(P), C.c_int(ishot), C.byref(g))
    if rc != 0:
        raise ValueError(f"bad geometry for shot {ishot} (rc={rc})")
    return g


def coef_table() -> np.ndarray:
    out = np.zeros(100, np.float32)
    lib().orc_coef_table(_fp(out))
    return out


def ricker(nt, fdom, dt):
    out = np.zeros(nt, np.float32)
    lib().orc_ricker(_fp(out), C.c_int(nt), C.c_float(fdom), C.c_float(dt))
    return out


def taper_row(nxl, nwin):
    out = np.zeros(nxl, np.float32)
    lib().orc_taper_row(_fp(out), C.c_int(nxl), C.c_int(nwin))
    return out


def window_pad(model, nx1, nx2, npml, square):
    nz, nx = model.shape
    nxl = nx2 - nx1 + 1
    out = np.zeros((nz + 2 * npml, nxl + 2 * npml), np.float32)
    m = np.ascontiguousarray(model, np.float32)
    lib().orc_window_pad(_fp(out), _fp(m), C.c_int(nx), C.c_int(nz), C.c_int(nx1), C.c_int(nx2),
                         C.c_int(npml), C.c_int(1 if square else 0))
    return out


def mute_cuts(P, g, vp0):
    cut = np.zeros(max(g.nrec, 1), np.int32)
    v = np.ascontiguousarray(vp0, np.float32)
    lib().orc_mute_cuts(C.byref(P), C.byref(g), None, _fp(v), None, _fp(cut))
    return cut[:g.nrec]


def laplacian(image, dx, dz):
    img = np.ascontiguousarray(image, np.float32).copy()
    nz, nx = img.shape
    lib().orc_laplacian(_fp(img), C.c_int(nx), C.c_int(nz), C.c_int(int(dx)), C.c_int(int(dz)))
    return img


def normalize_multi(sg, ss):
    sg = np.ascontiguousarray(sg, np.float32)
    ss = np.ascontiguousarray(ss, np.float32)
    out = np.zeros_like(sg)
    lib().orc_normalize_multi(_fp(out), _fp(sg), _fp(ss), C.c_size_t(sg.size))
    return out


def scaleimage(illum, grad):
    il = np.ascontiguousarray(illum, np.float32)
    gr = np.ascontiguousarray(grad, np.float32).copy()
    lib().orc_scaleimage(_fp(il), _fp(gr), C.c_size_t(il.size))
    return gr


@dataclass
class ShotResult:
    seis: np.ndarray          # (nt, nrec) before mute
    g1: np.ndarray            # (nz, nxl) cross-correlation of the shot
    illum: np.ndarray         # (nz, nxl) source illumination
    snap_fwd: np.ndarray | None
    snap_rec: np.ndarray | None
    geom: Geom


def _variant(f64, korder):
    """Suffix of the arithmetic form: f32 reference order, k32 kernel order, f64."""
    if f64 and korder:
        raise ValueError("the kernel-order form is FP32 only")
    return "f64" if f64 else ("k32" if korder else "f32")


def run_shot(P, ishot, vp0, rho0, coef=None, *, f64=False, korder=False, literal=False, forward_only=False,
             nt_limit=0, snap_it=-1, seis_in=None, exact_reverse=False, extra_ic=False,
             snap_every=0, store_fmt=0, snap_fmt=0) -> ShotResult:
    g = geometry(P, ishot)
    coef = coef_table() if coef is None else np.ascontiguousarray(coef, np.float32)
    rt = np.float64 if f64 else np.float32
    seis = np.zeros((P.nt, max(g.nrec, 1)), np.float32)
    g1 = np.zeros((P.nz, g.nxl), rt)
    il = np.zeros((P.nz, g.nxl), rt)
    sf = np.zeros((g.nzp, g.nxp), np.float32) if snap_it >= 0 else None
    sr = np.zeros((g.nzp, g.nxp), np.float32) if snap_it >= 0 else None
    ic = np.zeros((4, P.nz, g.nxl), np.float32) if extra_ic else None
    o = ShotOpts(literal=int(literal), forward_only=int(forward_only), nt_limit=int(nt_limit),
                 snap_it=int(snap_it), vvpad=None, rhopad=None, exact_reverse=int(exact_reverse), ic_out=_fp(ic),
                 snap_every=int(snap_every), store_fmt=int(store_fmt), snap_fmt=int(snap_fmt))
    vp0 = np.ascontiguousarray(vp0, np.float32)
    rho0 = np.ascontiguousarray(rho0, np.float32)
    si = None if seis_in is None else np.ascontiguousarray(seis_in, np.float32)
    fn = getattr(lib(), "orc_shot_" + _variant(f64, korder))
    rc = fn(C.byref(P), C.c_int(ishot), _fp(vp0), _fp(rho0), _fp(coef), C.byref(o), _fp(seis), _fp(si),
            _fp(g1), _fp(il), _fp(sf), _fp(sr))
    if rc != 0:
        raise RuntimeError(f"oracle shot failed rc={rc}")
    res = ShotResult(seis[:, :g.nrec], g1, il, sf, sr, g)
    if ic is not None:
        res.ic = {"g2": ic[0], "g3": ic[1], "g4": ic[2], "g3_true": ic[3]}
    return res


def store16_block(values, fmt):
    """The library's 16-bit boundary store on one strip block: returns (read
    back values, 16-bit codes, scale); fmt 1 binary16, 2 bfloat16."""
    v = np.array(values, np.float32, copy=True)
    codes = np.zeros(v.size, np.uint16)
    f = lib().orc_store16_block
    f.restype = C.c_float
    sc = f(_fp(v), C.c_longlong(v.size), C.c_int(int(fmt)), _fp(codes))
    return v, codes, float(sc)


def ic_cells(ps, vxs, vzs, p, vxr, vzr, ps_pre, p_pre, vv, npml, dx, dz, dt, korder=True):
    """One step of the decomposition / Poynting / energy-norm imaging conditions on
    given fields (reference shapes); returns the accumulators ud, du, lr, rl, g4
    (nzi, nxl) and g31, g32 (61, nzi, nxl), starting from zero."""
    nzp, nxp = ps.shape
    nzi, nxl = nzp - 2 * npml, nxp - 2 * npml
    f = lambda a: np.ascontiguousarray(a, np.float32)
    icb = np.zeros((5 + 2 * 61, nzi, nxl), np.float32)
    args = [f(a) for a in (ps, vxs, vzs, p, vxr, vzr, ps_pre, p_pre, vv)]
    getattr(lib(), "orc_ic_cells_" + _variant(False, korder))(C.c_int(nxp), C.c_int(nzp), C.c_int(npml),
                                                             *[_fp(a) for a in args], C.c_float(dx), C.c_float(dz),
                                                             C.c_float(dt), _fp(icb))
    return dict(ud=icb[0], du=icb[1], lr=icb[2], rl=icb[3], g4=icb[4], g31=icb[5:66], g32=icb[66:127])


def fd_steps(P, ishot, vp0, rho0, fields_, nsteps, *, it0=0, with_source=False, literal=False,
             f64=False, korder=False, coef=None):
    """Advance given wavefields nsteps forward steps in place.
    fields_: dict with p, vx, vz, pl1, pl2, pr1, pr2, pt1, pt2, pb1, pb2 (reference shapes)."""
    coef = coef_table() if coef is None else np.ascontiguousarray(coef, np.float32)
    fn = getattr(lib(), "orc_fd_steps_" + _variant(f64, korder))
    names = ["p", "vx", "vz", "pl1", "pl2", "pr1", "pr2", "pt1", "pt2", "pb1", "pb2"]
    rt = np.float64 if f64 else np.float32
    for n in names:
        assert fields_[n].dtype == rt and fields_[n].flags.c_contiguous
    rc = fn(C.byref(P), C.c_int(ishot), _fp(np.ascontiguousarray(vp0, np.float32)),
            _fp(np.ascontiguousarray(rho0, np.float32)), _fp(coef), C.c_int(int(literal)),
            C.c_int(nsteps), C.c_int(it0), C.c_int(int(with_source)), *[_fp(fields_[n]) for n in names])
    if rc != 0:
        raise RuntimeError(f"oracle fd_steps failed rc={rc}")


def alloc_fields(g, npml, f64=False):
    rt = np.float64 if f64 else np.float32
    nxp, nzp, nxl = g.nxp, g.nzp, g.nxl
    return dict(p=np.zeros((nzp, nxp), rt), vx=np.zeros((nzp, nxp - 1), rt), vz=np.zeros((nzp - 1, nxp), rt),
                pl1=np.zeros((nzp, npml), rt), pl2=np.zeros((nzp, npml), rt),
                pr1=np.zeros((nzp, npml), rt), pr2=np.zeros((nzp, npml), rt),
                pt1=np.zeros((npml, nxl), rt), pt2=np.zeros((npml, nxl), rt),
                pb1=np.zeros((npml, nxl), rt), pb2=np.zeros((npml, nxl), rt))


def rtm(P, vp0, rho0, coef=None, *, literal=False, stack64=True, f64=False, korder=False, snap_every=0,
        store_fmt=0, snap_fmt=0):
    """Whole rtm_model() job: returns (image1, image2s, illum) as float32 (nz, nx).
    snap_every > 0: with the library's decimated snapshots (exact values) restored
    into the reconstructed source field, which is not the reference's semantics.
    store_fmt / snap_fmt 1 / 2: the library's binary16 / bfloat16 boundary store and
    snapshots (FP32 forms)."""
    coef = coef_table() if coef is None else np.ascontiguousarray(coef, np.float32)
    img1 = np.zeros((P.nz, P.nx), np.float32)
    img2 = np.zeros((P.nz, P.nx), np.float32)
    ill = np.zeros((P.nz, P.nx), np.float32)
    done = C.c_int(0)
    fn = getattr(lib(), "orc_rtm_" + _variant(f64, korder))
    rc = fn(C.byref(P), _fp(np.ascontiguousarray(vp0, np.float32)), _fp(np.ascontiguousarray(rho0, np.float32)),
            _fp(coef), C.c_int(int(literal)), C.c_int(1 if (stack64 or f64) else 0), C.c_int(int(snap_every)),
            C.c_int(int(store_fmt)), C.c_int(int(snap_fmt)), _fp(img1), _fp(img2), _fp(ill), C.byref(done))
    if rc != 0:
        raise RuntimeError(f"oracle rtm failed rc={rc}")
    return img1, img2, ill
