"""gubingluo_amd -- MI355X-native 2-D acoustic prestack RTM (drop-in for gubingluo's rtm_gpu).

The compute path is librtm_hip.so (hand-written gfx950 HIP kernels behind a
C ABI, include/rtm_hip.h).  This module is a thin ctypes binding of that ABI
for tests, the bench and Python callers; it holds no compute of its own and
raises if the native library is not built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "librtm_hip.so")
EXE_PATH = os.path.join(PKG, "rtm_gpu")

STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2

# every symbol include/rtm_hip.h declares
EXPORTED = [
    "rtm_model", "rtm_real", "rtmh_last_error", "rtmh_version", "rtmh_device_count", "rtmh_create",
    "rtmh_destroy", "rtmh_shot_geometry", "rtmh_run", "rtmh_model_create", "rtmh_model_destroy",
    "rtmh_run_model", "rtmh_fd_steps", "rtmh_copy_probe", "rtmh_coef_table", "rtmh_read_parfile",
    "rtmh_get_constant", "rtmh_interp_model", "rtmh_selfdoc", "rtmh_postprocess", "rtmh_driver_main",
    "rtmh_server_create", "rtmh_server_submit", "rtmh_server_wait", "rtmh_server_query", "rtmh_server_destroy",
    "rtmh_scan_gather", "rtmh_read_gather", "rtmh_interp_seis", "rtmh_gather_geometry",
    "rtmh_server_request_bytes", "rtmh_smooth_gpu", "rtmh_postprocess_noalloc",
]


class Params(C.Structure):
    """rtmh_params: the scalar arguments of rtm_model() (rtm_model.cu:12-17), same order."""
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
    ]


class Options(C.Structure):
    _fields_ = [("batch", C.c_int), ("store_fmt", C.c_int), ("graph_chunk", C.c_int), ("tile_z", C.c_int),
                ("write_files", C.c_int), ("verbose", C.c_int), ("mem_budget", C.c_double), ("profile", C.c_int),
                ("shot_offsetmax", C.c_void_p), ("checkpoint", C.c_char_p), ("resume", C.c_int),
                ("ic_extra", C.c_int), ("shot_field", C.c_void_p), ("seis_wait", C.c_void_p),
                ("seis_user", C.c_void_p), ("snap_every", C.c_int), ("snap_fmt", C.c_int)]


# the other imaging conditions (rtmh_outputs order)
IC_IMAGES = ("image3s", "image4s", "image5s", "image3m", "image4m", "image5m", "imagetrue")


class Outputs(C.Structure):
    _fields_ = [("image1", C.c_void_p), ("image2s", C.c_void_p), ("illum", C.c_void_p),
                ("image_lap", C.c_void_p), ("image_norm", C.c_void_p),
                ("shot_seis", C.c_void_p), ("shot_g1", C.c_void_p), ("shot_illum", C.c_void_p),
                ("shot_seis_in", C.c_void_p)] + [(n, C.c_void_p) for n in IC_IMAGES]


class Stats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("gpu_ms", C.c_double), ("fwd_ms", C.c_double), ("rev_ms", C.c_double),
                ("cell_updates", C.c_double), ("nshots", C.c_int), ("batches", C.c_int), ("tile_z", C.c_int),
                ("batch", C.c_int), ("hbm_bytes_alg", C.c_double), ("kern_ms", C.c_double * 4),
                ("kern_calls", C.c_int * 4), ("step_bytes_fwd", C.c_double), ("step_bytes_rev", C.c_double),
                ("tile_z_rev", C.c_int), ("_pad0", C.c_int), ("fwd_bytes_alg", C.c_double),
                ("rev_bytes_alg", C.c_double)]


class Geom(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ishot", "nx1", "nx2", "nxl", "nxp", "nzp", "nsx", "nsz", "nrec",
                                       "rcol0", "noffset")]


class Grid(C.Structure):
    _fields_ = [("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float), ("nx", C.c_int), ("nz", C.c_int),
                ("nt", C.c_int), ("lx", C.c_int), ("lz", C.c_int), ("vpmin", C.c_float), ("vpmax", C.c_float)]


class Parfile(C.Structure):
    _fields_ = [("dt0", C.c_float), ("nt0", C.c_int), ("tsnap", C.c_float),
                ("xmax", C.c_float), ("zmax", C.c_float), ("dx0", C.c_float), ("dz0", C.c_float),
                ("fdom", C.c_float), ("amp", C.c_float), ("alp", C.c_float),
                ("spx0", C.c_float), ("spz0", C.c_float), ("dspx", C.c_float),
                ("offset", C.c_float), ("offsetmax", C.c_float),
                ("startshot", C.c_int), ("nshot", C.c_int), ("dshot", C.c_int), ("nshotall", C.c_int),
                ("npml", C.c_int), ("medium_flag", C.c_int), ("ismth", C.c_int),
                ("direct", C.c_float), ("direct0", C.c_float),
                ("idevice", C.c_int), ("flag_layout", C.c_int), ("mode", C.c_int), ("dataflag", C.c_int),
                ("sgyflag", C.c_int), ("endianflag", C.c_int), ("rhoflag", C.c_int),
                ("vp_file", C.c_char * 256), ("shotgather", C.c_char * 256), ("rho_file", C.c_char * 256)]


_lib = None


def lib() -> C.CDLL:
    """The native library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"librtm_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.rtmh_last_error.restype = C.c_char_p
        L.rtmh_version.restype = C.c_char_p
        _lib = L
    return _lib


def last_error() -> str:
    return lib().rtmh_last_error().decode()


def device_count() -> int:
    return int(lib().rtmh_device_count())


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_params(**kw) -> Params:
    d = dict(nshot=1, startshot=1, dshot=1, mode=0, medium_flag=0, flag_layout=2, sgyflag=0, endianflag=0,
             ismth=0, nx=100, nz=60, nt=200, dx=10.0, dz=10.0, dt=1e-3, npml=32, direct=1.5, direct0=0.05,
             tsnap=0.1, fdom=20.0, spx0=500.0, spz0=20.0, dspx=50.0, offset=0.0, offsetmax=1000.0, amp=1.0,
             alp=0.05)
    d.update(kw)
    P = Params()
    for k, v in d.items():
        setattr(P, k, v)
    return P


def coef_table() -> np.ndarray:
    out = np.zeros(100, np.float32)
    lib().rtmh_coef_table(_fp(out))
    return out


def geometry(P: Params, ishot: int) -> Geom:
    g = Geom()
    if lib().rtmh_shot_geometry(C.byref(P), C.c_int(ishot), C.byref(g)) != 0:
        raise ValueError(f"shot {ishot} outside the model")
    return g


def gather_geometry(P: Params, ishot: int, field) -> Geom:
    """Window of a recorded gather, field = (sx, offset, offsetmax) (rtm_real.cu:121-136)."""
    g = Geom()
    f = np.ascontiguousarray(field, np.float32)
    if lib().rtmh_gather_geometry(C.byref(P), C.c_int(ishot), _fp(f), C.byref(g)) != 0:
        raise ValueError(f"gather {ishot} outside the model")
    return g


def scan_gather(path: str, sgyflag: int = 0, max_shots: int = 4096):
    """Shot table of a gather file (scanshotgather): (table[4][nshots], endianflag)."""
    prof = np.zeros((4, max_shots), np.int32)
    e, n = C.c_int(0), C.c_int(0)
    rc = lib().rtmh_scan_gather(os.fsencode(path), C.c_int(sgyflag), C.byref(e), _fp(prof), C.c_int(max_shots),
                                C.byref(n))
    if rc != 0:
        raise RuntimeError(f"rtmh_scan_gather failed ({rc}): {last_error()}")
    return prof[:, :n.value].copy(), e.value


def read_gather(path: str, itraceall: int, ntrace: int, nt0: int, ncols: int, sgyflag: int = 0,
                endianflag: int = 0) -> np.ndarray:
    out = np.zeros((nt0, ncols), np.float32)
    rc = lib().rtmh_read_gather(os.fsencode(path), C.c_int(sgyflag), C.c_int(endianflag), C.c_int(itraceall),
                                C.c_int(ntrace), C.c_int(nt0), _fp(out), C.c_int(ncols))
    if rc != 0:
        raise RuntimeError(f"rtmh_read_gather failed ({rc}): {last_error()}")
    return out


SEIS_WAIT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int)


def interp_seis(seisobs, nx: int, nt: int, dt: float, dt0: float, out=None) -> np.ndarray:
    """Interpseis2d: seisobs[nt0][nx0] at dt0 onto [nt][nx] at dt (into `out` when given)."""
    a = np.ascontiguousarray(seisobs, np.float32)
    nt0, nx0 = a.shape
    if out is None:
        out = np.zeros((nt, nx), np.float32)
    assert out.shape == (nt, nx) and out.dtype == np.float32 and out.flags.c_contiguous
    rc = lib().rtmh_interp_seis(_fp(out), _fp(a), C.c_int(nx), C.c_int(nx0), C.c_int(nt), C.c_int(nt0),
                                C.c_float(dt), C.c_float(dt0))
    if rc != 0:
        raise ValueError(f"rtmh_interp_seis failed ({rc}): {last_error()}")
    return out


def shot_list(P: Params):
    return list(range(P.startshot, P.nshot + 1, P.dshot))


@dataclass
class RunResult:
    image1: np.ndarray
    image2s: np.ndarray
    illum: np.ndarray
    image_lap: np.ndarray | None
    image_norm: np.ndarray | None
    shot_seis: list | None
    shot_g1: list | None
    shot_illum: list | None
    stats: Stats
    ic: dict | None = None    # ic_extra: image3s, image4s, image5s, image3m, image4m, image5m, imagetrue


class Context:
    """One GPU's engine (rtmh_ctx)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        if lib().rtmh_create(C.c_int(device), C.byref(self._h)) != 0:
            raise RuntimeError(f"rtmh_create failed: {last_error()}")
        self.device = device

    def close(self):
        if self._h:
            lib().rtmh_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def model_create(self, vp0, rho0) -> int:
        """Upload a velocity/density model to HBM; returns its id."""
        vp0 = np.ascontiguousarray(vp0, np.float32)
        rho0 = np.ascontiguousarray(rho0, np.float32)
        assert vp0.shape == rho0.shape and vp0.ndim == 2
        mid = C.c_int(-1)
        nz, nx = vp0.shape
        if lib().rtmh_model_create(self._h, C.c_int(nx), C.c_int(nz), _fp(vp0), _fp(rho0), C.byref(mid)) != 0:
            raise RuntimeError(f"rtmh_model_create failed: {last_error()}")
        return int(mid.value)

    def model_destroy(self, mid: int) -> None:
        lib().rtmh_model_destroy(self._h, C.c_int(mid))

    def run(self, P: Params, vp0=None, rho0=None, coef=None, *, model=None, batch=0, store_fmt=STORE_F32,
            graph_chunk=0, tile_z=0, write_files=False, verbose=False, mem_budget=0.0, want_shots=False,
            want_post=True, seis_in=None, profile=False, shot_offsetmax=None, checkpoint=None,
            resume=False, ic_extra=False, shot_field=None, seis_wait=None, snap_every=0,
            snap_fmt=STORE_F16) -> RunResult:
        """One rtm_model() job; pass vp0/rho0 arrays, or model=<id> of a resident model.
        checkpoint: path written after every batch (FP64 stacks, shots done); resume=True
        continues a job from it.  shot_field: per shot (sx, offset, offsetmax) of recorded
        gathers (the geometry of rtm_real).  seis_wait: a SEIS_WAIT callback (user, shot index)
        the engine calls before it reads shot k of seis_in, after the forward pass.
        snap_every: decimated source-wavefield snapshots every that many steps
        (snap_fmt precision), restored in the reverse pass."""
        coef = coef_table() if coef is None else np.ascontiguousarray(coef, np.float32)
        nz, nx = P.nz, P.nx
        if model is None:
            vp0 = np.ascontiguousarray(vp0, np.float32)
            rho0 = np.ascontiguousarray(rho0, np.float32)
            assert vp0.shape == (nz, nx) and rho0.shape == (nz, nx)
        img1 = np.zeros((nz, nx), np.float32)
        img2 = np.zeros((nz, nx), np.float32)
        ill = np.zeros((nz, nx), np.float32)
        lap = np.zeros((nz, nx), np.float32) if want_post else None
        nrm = np.zeros((nz, nx), np.float32) if want_post else None
        out = Outputs(_fp(img1), _fp(img2), _fp(ill), _fp(lap), _fp(nrm), None, None, None, None)
        ic = {n: np.zeros((nz, nx), np.float32) for n in IC_IMAGES} if ic_extra else None
        if ic is not None:
            for n, a in ic.items():
                setattr(out, n, _fp(a))
        shots = shot_list(P)
        keep = []
        seis = g1s = ils = None
        if want_shots or seis_in is not None:
            geoms = []
            for k, s in enumerate(shots):
                Pk = P
                if shot_offsetmax is not None:
                    Pk = Params.from_buffer_copy(P)
                    Pk.offsetmax = float(shot_offsetmax[k])
                geoms.append(geometry(Pk, s) if shot_field is None else gather_geometry(P, s, shot_field[k]))
        if want_shots:
            seis = [np.zeros((P.nt, g.nrec), np.float32) for g in geoms]
            g1s = [np.zeros((nz, g.nxl), np.float32) for g in geoms]
            ils = [np.zeros((nz, g.nxl), np.float32) for g in geoms]
            arrs = []
            for lst in (seis, g1s, ils):
                pa = (C.c_void_p * len(lst))(*[_fp(a) for a in lst])
                arrs.append(pa)
            keep.extend(arrs)
            out.shot_seis = C.cast(arrs[0], C.c_void_p)
            out.shot_g1 = C.cast(arrs[1], C.c_void_p)
            out.shot_illum = C.cast(arrs[2], C.c_void_p)
        if seis_in is not None:
            si = [np.ascontiguousarray(a, np.float32) for a in seis_in]
            for a, g in zip(si, geoms):
                assert a.shape == (P.nt, g.nrec)
            pa = (C.c_void_p * len(si))(*[_fp(a) for a in si])
            keep.extend([pa, si])
            out.shot_seis_in = C.cast(pa, C.c_void_p)
        som = None if shot_offsetmax is None else np.ascontiguousarray(shot_offsetmax, np.float32)
        keep.append(som)
        sf = None if shot_field is None else np.ascontiguousarray(shot_field, np.float32).reshape(-1, 3)
        keep.append(sf)
        ck = None if checkpoint is None else os.fsencode(checkpoint)
        opt = Options(batch, store_fmt, graph_chunk, tile_z, int(write_files), int(verbose), float(mem_budget),
                      int(profile), _fp(som), ck, int(resume), int(ic_extra), _fp(sf))
        if seis_wait is not None:
            keep.append(seis_wait)
            opt.seis_wait = C.cast(seis_wait, C.c_void_p)
        opt.snap_every, opt.snap_fmt = int(snap_every), int(snap_fmt)
        st = Stats()
        if model is None:
            rc = lib().rtmh_run(self._h, C.byref(P), _fp(vp0), _fp(rho0), _fp(coef), C.byref(opt), C.byref(out),
                                C.byref(st))
        else:
            rc = lib().rtmh_run_model(self._h, C.c_int(model), C.byref(P), _fp(coef), C.byref(opt), C.byref(out),
                                      C.byref(st))
        if rc != 0:
            raise RuntimeError(f"rtmh_run failed ({rc}): {last_error()}")
        return RunResult(img1, img2, ill, lap, nrm, seis, g1s, ils, st, ic)

    def smooth(self, window, npml: int = 0, nsp: int = 0) -> np.ndarray:
        """pmlvelsmooth1d (npml > 0) then velsmooth1d (nsp > 0) of a window on the GPU."""
        w = np.array(window, np.float32, order="C", copy=True)
        nz, nx = w.shape
        if lib().rtmh_smooth_gpu(self._h, _fp(w), C.c_int(nx), C.c_int(nz), C.c_int(npml), C.c_int(nsp)) != 0:
            raise RuntimeError(f"rtmh_smooth_gpu failed: {last_error()}")
        return w

    def copy_probe(self, nbytes: int, reps: int = 3) -> float:
        """Device-to-device copy bandwidth, GB/s counting read + written bytes."""
        g = C.c_double(0.0)
        if lib().rtmh_copy_probe(self._h, C.c_double(float(nbytes)), C.c_int(reps), C.byref(g)) != 0:
            raise RuntimeError(f"rtmh_copy_probe failed: {last_error()}")
        return float(g.value)

    def fd_steps(self, P: Params, ishot, vp0, rho0, fields, nsteps, *, it0=0, with_source=False, coef=None):
        """Advance fields (dict p, vx, vz, p1, p2 of shape (nzp, nxp), float32) in place."""
        coef = coef_table() if coef is None else np.ascontiguousarray(coef, np.float32)
        for k in ("p", "vx", "vz", "p1", "p2"):
            assert fields[k].dtype == np.float32 and fields[k].flags.c_contiguous
        rc = lib().rtmh_fd_steps(self._h, C.byref(P), C.c_int(ishot), _fp(np.ascontiguousarray(vp0, np.float32)),
                                 _fp(np.ascontiguousarray(rho0, np.float32)), _fp(coef), C.c_int(nsteps),
                                 C.c_int(it0), C.c_int(int(with_source)), _fp(fields["p"]), _fp(fields["vx"]),
                                 _fp(fields["vz"]), _fp(fields["p1"]), _fp(fields["p2"]))
        if rc != 0:
            raise RuntimeError(f"rtmh_fd_steps failed ({rc}): {last_error()}")


class ServerOptions(C.Structure):
    _fields_ = [("mem_budget", C.c_double), ("max_slots", C.c_int), ("tile_z", C.c_int), ("store_fmt", C.c_int)]


class RequestStats(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double),
                ("cell_updates", C.c_double), ("nshots", C.c_int), ("status", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("ticks", C.c_longlong), ("launches", C.c_longlong), ("peak_slots", C.c_int),
                ("active_slots", C.c_int), ("queued_requests", C.c_int), ("busy_s", C.c_double),
                ("cell_updates", C.c_double), ("budget_bytes", C.c_double), ("allocations", C.c_longlong),
                ("alloc_s", C.c_double)]


class Server:
    """Continuous-batching RTM server on one context (rtmh_server_*)."""

    def __init__(self, ctx: Context, *, mem_budget=0.0, max_slots=0, tile_z=0, store_fmt=STORE_F32):
        self._h = C.c_void_p()
        o = ServerOptions(float(mem_budget), int(max_slots), int(tile_z), int(store_fmt))
        if lib().rtmh_server_create(ctx._h, C.byref(o), C.byref(self._h)) != 0:
            raise RuntimeError(f"rtmh_server_create failed: {last_error()}")
        self.ctx = ctx

    def submit(self, model: int, P: Params, coef=None) -> int:
        coef = coef_table() if coef is None else np.ascontiguousarray(coef, np.float32)
        rid = C.c_int(-1)
        if lib().rtmh_server_submit(self._h, C.c_int(model), C.byref(P), _fp(coef), C.byref(rid)) != 0:
            raise RuntimeError(f"rtmh_server_submit failed: {last_error()}")
        return int(rid.value)

    def request_bytes(self, model: int, P: Params, coef=None):
        """(bytes of the job's largest shot while resident, bytes of its request block)."""
        coef = coef_table() if coef is None else np.ascontiguousarray(coef, np.float32)
        a, b = C.c_double(0.0), C.c_double(0.0)
        if lib().rtmh_server_request_bytes(self._h, C.c_int(model), C.byref(P), _fp(coef), C.byref(a),
                                           C.byref(b)) != 0:
            raise RuntimeError(f"rtmh_server_request_bytes failed: {last_error()}")
        return float(a.value), float(b.value)

    def wait(self, rid: int, P: Params, timeout=0.0):
        """Block until request rid is done; returns (image1, image2s, illum, image_lap, image_norm, stats)."""
        nz, nx = P.nz, P.nx
        ims = [np.zeros((nz, nx), np.float32) for _ in range(5)]
        out = Outputs(*[_fp(a) for a in ims], None, None, None, None)
        st = RequestStats()
        rc = lib().rtmh_server_wait(self._h, C.c_int(rid), C.c_double(timeout), C.byref(out), C.byref(st))
        if rc == 1:
            raise TimeoutError(f"request {rid} not done")
        if rc != 0:
            raise RuntimeError(f"rtmh_server_wait failed ({rc}): {last_error()}")
        return (*ims, st)

    def query(self) -> ServerStats:
        st = ServerStats()
        lib().rtmh_server_query(self._h, C.byref(st))
        return st

    def close(self):
        if self._h:
            lib().rtmh_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
