"""The benchmark configurations at their stated sizes, through the C ABI on
the GPU.  Each checks what can be checked at that size: the kernel-order
oracle bit for bit where the CPU affords it, the reference-order oracle
within the parity bars (seismogram 1e-5, image 1e-4), and otherwise
size-independent properties (batch, graph, tile-height and mix invariance;
reduced-precision stores within 1e-3 of the FP32 store).

  single-shot anomaly   400x200, npml 40, nt 2000, 400 receivers
  batched ragged        2048x768, 64 shots with windows of 400-1600 columns
                        in one launch batch, graph chunk 50, autotuned tiles
  precision variant     2048x1024 dx 7.5 m, nt 6000, FP32/FP16/BF16 stores
  serving               a Poisson stream over 8 resident models, grids up to
                        2048x1024, FP16 boundaries
"""
import time

import numpy as np
import pytest

import gubingluo_amd as G
from gubingluo_amd import synth
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _odict(P):
    return O.make_params(**{n: getattr(P, n) for n, _ in P._fields_})


def test_single_shot_anomaly_full_size(gpu_ctx):
    """400x200 two-layer model with the 3000 m/s anomaly, npml 40, Ricker
    20 Hz, dt 1 ms, nt 2000, 400 surface receivers (rtm_model.cu:626-652)."""
    P, vp, rho = synth.single_shot_anomaly()
    Po = _odict(P)
    print(f"[cfg single-shot] {P.nx}x{P.nz} nt={P.nt} npml={P.npml}", flush=True)
    r = gpu_ctx.run(P, vp, rho, want_shots=True)
    k = O.run_shot(Po, 1, vp, rho, korder=True)
    o = O.run_shot(Po, 1, vp, rho)
    e = (rel_l2(r.shot_seis[0], o.seis), rel_l2(r.shot_g1[0], o.g1), rel_l2(r.shot_illum[0], o.illum))
    print(f"[cfg single-shot] vs reference order: seis {e[0]:.2e} g1 {e[1]:.2e} illum {e[2]:.2e}", flush=True)
    assert e[0] <= 1e-5 and e[1] <= 1e-4 and e[2] <= 1e-4
    assert np.array_equal(r.shot_seis[0], k.seis)
    assert np.array_equal(r.shot_g1[0], k.g1)
    assert np.array_equal(r.shot_illum[0], k.illum)
    i1, i2, il = O.rtm(Po, vp, rho, korder=True)
    assert np.array_equal(r.image1, i1) and np.array_equal(r.image2s, i2) and np.array_equal(r.illum, il)


def test_batched_ragged_64_shot_launch(gpu_ctx):
    """One 64-shot batch of the 2048x768 survey with windows drawn from
    400-1600 columns: graph chunk 50 and autotuned tile heights equal one shot
    per launch with 16-row tiles and plain launches, shot by shot; the
    narrowest window equals the kernel-order oracle."""
    P, vp, rho, om = synth.batched_ragged()
    P.startshot, P.nshot, P.dshot = 1, 64, 1
    sub = om[:64]
    mid = gpu_ctx.model_create(vp, rho)
    try:
        print(f"[cfg batched] {P.nx}x{P.nz} nt={P.nt} 64 shots, windows "
              f"{int(sub.min() / P.dx) + 1}..{int(sub.max() / P.dx) + 1}", flush=True)
        t0 = time.perf_counter()
        a = gpu_ctx.run(P, model=mid, batch=64, graph_chunk=50, tile_z=0, want_shots=True, shot_offsetmax=sub)
        print(f"[cfg batched] one launch batch: {time.perf_counter() - t0:.1f} s, tz {a.stats.tile_z}/"
              f"{a.stats.tile_z_rev}, batches {a.stats.batches}", flush=True)
        assert a.stats.batches == 1
        t0 = time.perf_counter()
        b = gpu_ctx.run(P, model=mid, batch=1, graph_chunk=0, tile_z=16, want_shots=True, shot_offsetmax=sub)
        print(f"[cfg batched] one shot per launch: {time.perf_counter() - t0:.1f} s", flush=True)
        for k in range(64):
            assert np.array_equal(a.shot_seis[k], b.shot_seis[k]), k
            assert np.array_equal(a.shot_g1[k], b.shot_g1[k]), k
            assert np.array_equal(a.shot_illum[k], b.shot_illum[k]), k
        assert np.array_equal(a.image1, b.image1) and np.array_equal(a.image2s, b.image2s)
        assert np.array_equal(a.illum, b.illum)
    finally:
        gpu_ctx.model_destroy(mid)
    # the narrowest window against the kernel-order oracle
    widths = [a.shot_g1[k].shape[1] for k in range(64)]
    k = int(np.argmin(widths))
    Po = _odict(P)
    Po.offsetmax = float(sub[k])
    print(f"[cfg batched] shot {k + 1} ({widths[k]} columns) vs oracle", flush=True)
    o = O.run_shot(Po, k + 1, vp, rho, korder=True)
    assert np.array_equal(a.shot_seis[k], o.seis)
    assert np.array_equal(a.shot_g1[k], o.g1)
    assert np.array_equal(a.shot_illum[k], o.illum)


def test_precision_variant_full_length(gpu_ctx):
    """2048x1024, dx 7.5 m, 25 Hz, dt 0.6 ms, nt 6000, one full-aperture
    shot: the FP16 and BF16 boundary stores stay within 1e-3 of the FP32
    store over the whole length."""
    P, vp, rho = synth.precision_grid()
    P.startshot = P.nshot = 64
    mid = gpu_ctx.model_create(vp, rho)
    try:
        res = {}
        for name, fmt in (("f32", G.STORE_F32), ("f16", G.STORE_F16), ("bf16", G.STORE_BF16)):
            t0 = time.perf_counter()
            res[name] = gpu_ctx.run(P, model=mid, store_fmt=fmt, graph_chunk=50, want_post=False)
            print(f"[cfg precision] {name}: {time.perf_counter() - t0:.1f} s", flush=True)
        ref = res["f32"]
        assert np.isfinite(ref.image1).all() and np.abs(ref.image1).max() > 0
        for name in ("f16", "bf16"):
            r = res[name]
            e = (rel_l2(r.image1, ref.image1), rel_l2(r.image2s, ref.image2s), rel_l2(r.illum, ref.illum))
            print(f"[cfg precision] {name} vs f32 store: image1 {e[0]:.2e} image2s {e[1]:.2e} illum {e[2]:.2e}",
                  flush=True)
            assert max(e) <= 1e-3, (name, e)
    finally:
        gpu_ctx.model_destroy(mid)


def test_precision_variant_vs_fp64_oracle(gpu_ctx):
    """The precision model family (dx 7.5 m, 25 Hz, dt 0.6 ms) on a shot the
    CPU affords in FP64 (401-column window, nt 2000): FP32 store within 1e-4
    of the FP64 oracle, FP16 and BF16 stores within 1e-3."""
    P, vp, rho = synth.precision_grid()
    P.startshot = P.nshot = 64
    P.flag_layout, P.offsetmax, P.nt = 1, 3000.0, 2000
    print("[cfg precision] FP64 oracle, 401-column window, nt 2000", flush=True)
    t0 = time.perf_counter()
    i1, i2, il = O.rtm(_odict(P), vp, rho, f64=True)
    print(f"[cfg precision] oracle {time.perf_counter() - t0:.1f} s", flush=True)
    for fmt, tol in ((G.STORE_F32, 1e-4), (G.STORE_F16, 1e-3), (G.STORE_BF16, 1e-3)):
        r = gpu_ctx.run(P, vp, rho, store_fmt=fmt, want_post=False)
        e = (rel_l2(r.image1, i1), rel_l2(r.image2s, i2), rel_l2(r.illum, il))
        print(f"[cfg precision] store {fmt} vs FP64: image1 {e[0]:.2e} image2s {e[1]:.2e} illum {e[2]:.2e}",
              flush=True)
        assert max(e) <= tol, (fmt, e)


def _serve_requests(max_gcell=300.0, min_req=4):
    models = synth.serve_models()
    reqs = synth.serve_stream(16, models)
    out, tot = [], 0.0
    for ta, m, P in reqs:
        nxp = min(P.nx, int(P.offsetmax / P.dx) + 1) + 2 * P.npml
        c = P.nshot * 3.0 * P.nt * nxp * (P.nz + 2 * P.npml) / 1e9
        if len(out) >= min_req and tot + c > max_gcell:
            break
        out.append((ta, m, P))
        tot += c
    return models, out, tot


def test_serving_stream_alone_mixed_and_engine(gpu_ctx):
    """A Poisson stream over the 8 resident models (grids 256^2..2048x1024,
    nt 1000-6000, 1-32 shots, FP16 boundaries): every request's images are
    the same served alone, served in the mix (arrivals as drawn), and run as
    a job through rtmh_run_model."""
    models, reqs, gcell = _serve_requests()
    mids = [gpu_ctx.model_create(v, r) for v, r in models]
    try:
        print(f"[cfg serve] {len(reqs)} requests, {gcell:.0f} Gcell, grids "
              f"{sorted({(P.nx, P.nz) for _, _, P in reqs})}", flush=True)
        eng = [gpu_ctx.run(P, model=mids[m], store_fmt=G.STORE_F16, batch=0, graph_chunk=50) for _, m, P in reqs]
        with G.Server(gpu_ctx, store_fmt=G.STORE_F16) as srv:
            for (ta, m, P), r in zip(reqs, eng):
                rid = srv.submit(mids[m], P)
                i1, i2, il, lap, nrm, st = srv.wait(rid, P, timeout=600)
                assert np.array_equal(i1, r.image1) and np.array_equal(i2, r.image2s)
                assert np.array_equal(il, r.illum) and np.array_equal(lap, r.image_lap)
                assert np.array_equal(nrm, r.image_norm)
            print("[cfg serve] alone == engine", flush=True)
            t0 = time.perf_counter()
            ids = []
            for ta, m, P in reqs:
                now = time.perf_counter() - t0
                if ta > now:
                    time.sleep(ta - now)
                ids.append(srv.submit(mids[m], P))
            spans = []
            for rid, (ta, m, P), r in zip(ids, reqs, eng):
                i1, i2, il, lap, nrm, st = srv.wait(rid, P, timeout=600)
                assert np.array_equal(i1, r.image1) and np.array_equal(i2, r.image2s)
                assert np.array_equal(il, r.illum)
                spans.append((st.start_s, st.done_s))
            q = srv.query()
            overlap = sum(1 for i in range(len(spans)) for j in range(i + 1, len(spans))
                          if spans[i][0] < spans[j][1] and spans[j][0] < spans[i][1])
            print(f"[cfg serve] mixed == engine; peak slots {q.peak_slots}, overlapping request pairs {overlap}",
                  flush=True)
            assert overlap >= 1                # requests shared the device
    finally:
        for m in mids:
            gpu_ctx.model_destroy(m)
