"""CPU restatement of the field-data path's host steps (test infrastructure
only: the checker of librtm_hip's rtmh_scan_gather / rtmh_read_gather /
rtmh_interp_seis and of rtm_real's geometry; never part of the product).

  scan_gather   scanshotgather           RTM_GPU_velstr.cpp:363-454
  read_gather   the gather read          rtm_real.cu:379-391
  kaiser21      Kaiser / modZeroBessel   Interp_function.cpp:10-40
  sinc_interp   Sincinterp               Interp_function.cpp:61-112
  fk_interp     GeneralFKinterp          Interp_function.cpp:121-243
  interp_seis2d Interpseis2d             Interp_function.cpp:245-264
  gather_apertures                       rtm_real.cu:121-136

FFTW's transforms are restated with numpy.fft (pocketfft): fftw FORWARD is
np.fft.fft* and BACKWARD is n * np.fft.ifft* (FFTW does not normalise).  The
f-k results therefore agree with any other FFT to rounding, not bit for bit.

fk_interp is pinned by closed forms of events constant over traces: for
lx = 1 the spectrum's path, for lx >= 2 the replication, the zero padding and
the block average of the padded spectrum (the box sums of the lx aliases give
every fine trace lx N(x mod lx) / ceil(nx0 / lx) times the lx = 1 output),
and the |H| > lx^2 clip (Interp_function.cpp:205-207) by two spikes whose
aliases cancel in Z: every |H| above lx^2 gives lx^2 times the lx = 1 output at
the spread traces, none above it gives the data plus a shifted copy, and two
amplitudes bracket the threshold.  Interp_function.cpp needs FFTW, which is
absent, so no output of the reference's own code is compared.
"""
from __future__ import annotations

import math

import numpy as np

TRACE_HDR = 240


def _hdr_fields(h: bytes, swap: bool):
    e = ">" if swap else "<"
    offset = int(np.frombuffer(h, e + "i4", 1, 36)[0])
    sx = int(np.frombuffer(h, e + "i4", 1, 72)[0])
    ns = int(np.frombuffer(h, e + "i2", 1, 114)[0])
    dt = int(np.frombuffer(h, e + "i2", 1, 116)[0])
    return offset, sx, ns, dt


def scan_gather(data: bytes, sgyflag: int = 0):
    """(table [4][nshots]: sx, traces, first trace (1-based), offset; endianflag).
    A new shot starts wherever the header's sx changes (RTM_GPU_velstr.cpp:409-453);
    headers are byte-swapped when ns or dt read as non-positive and both read
    positive swapped (:383-396)."""
    start = 3600 if sgyflag else 0
    _, _, ns, dt = _hdr_fields(data[start:start + TRACE_HDR], False)
    swap = False
    if dt <= 0 or ns <= 0:
        _, _, ns2, dt2 = _hdr_fields(data[start:start + TRACE_HDR], True)
        if dt2 > 0 and ns2 > 0:
            swap, ns = True, ns2
    tr = TRACE_HDR + 4 * ns
    ntr = (len(data) - start) // tr
    sxs, counts, firsts, offs = [], [], [], []
    for i in range(ntr):
        off, sx, ns_i, dt_i = _hdr_fields(data[start + i * tr:start + i * tr + TRACE_HDR], swap)
        if ns_i <= 0 or dt_i <= 0:
            break
        if not sxs or sx != sxs[-1]:
            sxs.append(sx), counts.append(0), firsts.append(i + 1), offs.append(off)
        counts[-1] += 1
    return np.array([sxs, counts, firsts, offs], np.int64), int(swap)


def read_gather(data: bytes, itraceall: int, ntrace: int, nt0: int, ncols: int, sgyflag=0, endianflag=0):
    """seis[nt0][ncols]: ntrace traces from trace itraceall (rtm_real.cu:379-391)."""
    start = (itraceall - 1) * (TRACE_HDR + 4 * nt0) + 3600 * (1 if sgyflag else 0)
    out = np.zeros((nt0, ncols), np.float32)
    dt = np.dtype(">f4" if endianflag else "<f4")
    for ix in range(min(ntrace, ncols)):
        o = start + ix * (TRACE_HDR + 4 * nt0) + TRACE_HDR
        out[:, ix] = np.frombuffer(data, dt, nt0, o).astype(np.float32)
    return out


def gather_apertures(sx: int, ntrace: int, off: int, dx0: float, dx: float):
    """rtm_real.cu:121-136 in float32: (field (sx, offset, offsetmax), ncol0, ncol)."""
    f32 = np.float32
    offsetmax = f32(off + (ntrace - 1) * dx0)
    offset, spx = f32(off), f32(sx)
    ncols = []
    for d in (f32(dx0), f32(dx)):
        noff = int(offset / d)
        n1 = int(spx / d)
        n2 = int(f32(spx + offsetmax) / d)
        ncols.append(n2 - n1 + 1 - noff)
    return (float(spx), float(offset), float(offsetmax)), ncols[0], ncols[1]


def bessel_i0_series(x: float) -> float:
    """modZeroBessel: 1 + sum_{i=1}^{19} ((x/2)^i / i!)^2"""
    h = x / 2
    num = fact = r = 1.0
    for i in range(1, 20):
        num *= h * h
        fact *= i
        r += num / (fact * fact)
    return r


def kaiser21(beta: float = 2.0) -> np.ndarray:
    m2 = 10.0
    den = bessel_i0_series(beta)
    w = np.zeros(21, np.float32)
    for n in range(21):
        v = (n - m2) / m2
        w[n] = np.float32(bessel_i0_series(beta * math.sqrt(1 - v * v)) / den)
    return w


def sinc_interp(seisin: np.ndarray, nx: int, nt: int, dt: float, dt0: float) -> np.ndarray:
    """Sincinterp: out[it][ix] = sum_k in[k][ix] * sinc((it*dt - k*dt0)/dt0) * w[k - s0 + w0]
    over the 21 input samples around (int)((it+1)*dt/dt0), float32 accumulation in k order."""
    f32 = np.float32
    nt0 = seisin.shape[0]
    a = np.ascontiguousarray(seisin[:, :nx], np.float32)
    dt, dt0 = f32(dt), f32(dt0)
    w = kaiser21()
    tn = np.arange(nt0, dtype=np.float32) * dt0
    out = np.zeros((nt, nx), np.float32)
    for it in range(nt):
        it0 = int(f32(f32(it + 1) * dt) / dt0)
        s0, s1 = max(0, it0 - 10), min(nt0 - 1, it0 + 10)
        ln = s1 - s0 + 1
        w0 = 21 - ln if s0 <= 10 else 0
        t = f32(f32(it) * dt)
        acc = np.zeros(nx, np.float32)
        for k in range(max(ln, 0)):
            x = f32((t - tn[k + s0]) / dt0)
            sc = f32(1.0) if x == 0 else f32(math.sin(math.pi * float(x)) / (math.pi * float(x)))
            acc = acc + (a[k + s0] * sc) * w[k + w0]
        out[it] = acc
    return out


def fk_interp(seisin: np.ndarray, nx: int, lx: int) -> np.ndarray:
    """GeneralFKinterp; seisin [nt0][nx0] -> [nt0][nx] (even nt0, nx <= lx*nx0)."""
    nt0, nx0 = seisin.shape
    assert nt0 % 2 == 0 and nx <= lx * nx0
    nf = 1 + nt0 // 2
    Nx, Nt = lx * nx0, lx * nt0
    B = np.ascontiguousarray(seisin.T, np.float64).astype(np.complex128)      # [nx0][nt0]
    K = np.fft.fft2(B)
    C = np.tile(K, (lx, 1))                                                  # [Nx][nt0]
    S = np.zeros((Nx, Nt), np.complex128)
    S[:nx0, :nt0] = B
    S = np.fft.fft2(S)
    Z = S.reshape(lx, nx0, Nt).sum(axis=0) / lx
    Z = np.tile(Z, (lx, 1))
    zmax = np.float32(np.max(np.sqrt((Z.real ** 2 + Z.imag ** 2).astype(np.float32))))
    Z = Z + float(zmax) * 0.00001
    H = S[:, :nf] / Z[:, :nf]
    big = np.sqrt((H.real ** 2 + H.imag ** 2).astype(np.float32)) > lx * lx
    H[big] = float(lx)
    G1 = np.fft.ifft(H * C[:, :nf], axis=0) * Nx                             # [Nx][nf], backward over k
    G = np.zeros((Nx, nt0), np.complex128)
    G[:, :nf] = G1
    it1 = (nt0 + 1) // 2
    for it in range(nf, nt0):
        G[:, it] = np.conj(G1[:, it1])
        it1 -= 1
    G = np.fft.ifft(G, axis=1) * nt0
    return np.ascontiguousarray(G[:nx].real.T).astype(np.float32)


def interp_seis2d(seisobs: np.ndarray, nx: int, nt: int, dt: float, dt0: float) -> np.ndarray:
    """Interpseis2d; where neither axis is refined the samples are copied
    (the reference leaves its record unset there)."""
    nt0, nx0 = seisobs.shape
    if nx > nx0:
        lx = (nx - 1) // (nx0 - 1)
        tmp = fk_interp(seisobs, nx, lx)
        return sinc_interp(tmp, nx, nt, dt, dt0) if nt > nt0 else tmp[:nt].copy()
    if nt > nt0:
        return sinc_interp(seisobs, nx, nt, dt, dt0)
    return np.ascontiguousarray(seisobs[:nt, :nx], np.float32)
