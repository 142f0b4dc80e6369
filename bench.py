#!/usr/bin/env python3
"""bench.py -- 2-D acoustic prestack RTM on one MI355X.

Default workload (the headline configuration, --workload survey): plain RTM
of a 200-shot survey over a seeded smooth-random layered model 1500-4500 m/s,
nx=1024 nz=512, dx=10 m, PML 50, Ricker 15 Hz, dt=0.8 ms nt=4000, shots every
50 m, split-spread 301 receivers (layout 1), fp32 stencil, all three imaging
conditions (cross-correlation, source-illumination normalised, Laplacian
filtered).  Data: modelled by the same code from the synthetic model.

A "step" is the migration of one batch of --shots-per-step shots (forward
modelling with boundary saving, reverse reconstruction + receiver
propagation + imaging, FP64 stacking); by default the whole 200-shot survey,
all shots in flight in one launch batch (about 110 GB of slots), so that
the per-step launches carry enough tiles to keep the tail short.  The velocity model is resident in
HBM before the timed region (rtmh_model_create); every step runs through the
C ABI (rtmh_run_model), which returns after its stream has drained.
--ic-extra adds the other imaging conditions that the rtm_model drop-in always
computes (decomposed, Poynting and energy-norm images).

The default run also measures, as sub-objects of the same line (skip with
--quick): "plain_form" (survey shots one per launch with per-step launches,
the plain GPU form), "single_shot" (the single-shot configuration, 400x200
nt 2000, on the GPU in the plain form and the default engine form; the CPU
baseline runs it too), "error_vs_oracle" (the GPU's images of the survey shots
the CPU baseline runs, against those of the oracle), "batched" (one 64-shot
launch batch of the ragged 2048x768
survey), "precision" (2048x1024 dx=7.5 m nt=6000 shots with FP32 / FP16 /
BF16 boundary stores, with FP16 / BF16 source-wavefield snapshots on top of
the 16-bit strips, and the image error of each form against the FP64 oracle
on a shot the CPU affords), "served" (a Poisson stream of --requests
requests against the plain one-shot-per-launch form) and, in
"cpu_baseline", the oracle on all host cores and on one core, on the
single-shot configuration (400x200, nt 2000) and on survey shots.

Other workloads (one JSON line each, same keys):
  --workload batched    2048x768, 256 shots with per-shot aperture windows
                        U[400,1600] columns, 64 shots per launch, hipGraph
                        time loop, autotuned tile height
  --workload precision  2048x1024 dx=7.5 m nt=6000 25 Hz: throughput with
                        FP32 / FP16 / BF16 boundary stores, and the image
                        error of each against the FP64 CPU oracle
  --workload serve      Poisson stream (4 req/s) over 8 resident models,
                        grids 256^2..2048x1024, nt 1000-6000, 1-32 shots,
                        FP16 boundaries, continuous batching; against the
                        plain one-shot-per-launch GPU form on the same stream
  --workload single     the single-shot configuration (400x200, npml 40,
                        nt 2000) on the GPU: plain form and engine form
  --workload field      field-data path on the survey model: 128 recorded
                        one-sided gathers (151 traces at 20 m, 2000 samples
                        at 1.6 ms, seeded random samples) scanned, read and
                        resampled to 301 traces at 10 m and 4000 samples at
                        0.8 ms on 16 host threads (f-k and sinc) beside the
                        GPU, migrated in batches of 32 (rtm_real's rule); the
                        step is timed whole (file to images)

Output: ONE JSON line on stdout (progress goes to stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cells_of(P, geoms):
    nt = P.nt
    return sum(2.0 * nt * g.nxp * g.nzp + (nt - 1.0) * P.nz * g.nxl for g in geoms)


def _omp_threads(n):
    """OpenMP threads of the oracle's parallel loops (0: the runtime default)."""
    import ctypes

    try:
        gomp = ctypes.CDLL("libgomp.so.1")
    except OSError:
        return
    gomp.omp_set_num_threads(int(n) if n > 0 else (os.cpu_count() or 1))


def _host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)


def cpu_shots(P_full, vp, rho, shots, min_seconds, threads, keep=None):
    """The oracle (FP32 restatement, OpenMP over rows) on whole shots of a job
    (forward + reconstruction + receiver + imaging at the full nt) until at
    least min_seconds of CPU work: (cells, seconds, shots run).  keep: a dict
    that receives each shot's (g1, illum)."""
    from oracle import oracle as O

    _omp_threads(threads)
    d = {n: getattr(P_full, n) for n, _ in P_full._fields_}
    cells, done = 0.0, []
    t0 = time.perf_counter()
    for ishot in shots:
        d.update(nshot=ishot, startshot=ishot)
        Po = O.make_params(**d)
        g = O.geometry(Po, ishot)
        r = O.run_shot(Po, ishot, vp, rho)
        if keep is not None:
            keep[ishot] = (r.g1, r.illum)
        cells += 2.0 * Po.nt * g.nxp * g.nzp + (Po.nt - 1.0) * Po.nz * g.nxl
        done.append(ishot)
        if time.perf_counter() - t0 >= min_seconds:
            break
    dt = time.perf_counter() - t0
    _omp_threads(0)
    return cells, dt, done


CPU_SHOTS = range(100, 106)    # survey shots the CPU baseline runs (as many as fit its time)


def cpu_baseline(P_full, vp, rho, synth, quick=False, keep=None):
    """The reported CPU baseline: the oracle on all host cores over whole
    survey shots (value), and on the single-shot configuration (the
    reference's CPU-runnable case) on all cores and on one core."""
    nth = _host_threads()
    cells, dt, done = cpu_shots(P_full, vp, rho, CPU_SHOTS, 10.0, nth, keep=keep)
    out = {"value": round(cells / dt / 1e9, 4), "unit": "Gcell/s", "cores": nth, "kind": "port",
           "sample": f"{len(done)} whole shot(s) #{done[0]}-#{done[-1]} of the 200-shot survey (nt={P_full.nt}): "
                     f"{cells / 1e9:.2f} Gcell in {dt:.1f} s; {len(done) / dt * 3600.0:.0f} shots/hour"}
    if quick:
        return out
    P1, v1, r1 = synth.single_shot_anomaly()
    c1, d1, _ = cpu_shots(P1, v1, r1, [1], 0.0, nth)
    cs, ds, _ = cpu_shots(P1, v1, r1, [1], 0.0, 1)
    out["single_shot_anomaly"] = {
        "config": "400x200, npml 40, nt 2000, 400 receivers, one whole shot",
        "all_cores": {"value": round(c1 / d1 / 1e9, 4), "unit": "Gcell/s", "cores": nth, "s_per_shot": round(d1, 2)},
        "one_core": {"value": round(cs / ds / 1e9, 4), "unit": "Gcell/s", "cores": 1, "s_per_shot": round(ds, 2)}}
    return out


def traffic_ratio(kernel):
    """HBM bytes / algorithmic bytes of a step kernel, measured offline with
    rocprofv3 FETCH_SIZE / WRITE_SIZE (profiles/step_kernel_traffic.json)."""
    try:
        with open(os.path.join(ROOT, "profiles", "step_kernel_traffic.json")) as f:
            return float(json.load(f)[kernel]["traffic_over_algorithmic"])
    except (OSError, KeyError, ValueError):
        return None


def roofline_of(recs, P, copy_gbs=None):
    """Dominant kernel: the reverse time step k_rev (one launch per step:
    reconstruction, receiver wavefield, imaging), timed by HIP events on the
    engine's stream around every reverse pass of the timed region; the
    algorithmic bytes come from the per-cell figures (engine stats
    rev_bytes_alg: all reverse steps of all batches), so a launch's average
    is exact also when the batches of a step differ in size."""
    rev_steps = sum(s.batches * (P.nt - 1) for s, _, _ in recs)
    rev_ms = sum(s.rev_ms for s, _, _ in recs)
    # totals over every batch of every step (batches may differ in size)
    step_bytes = sum(s.rev_bytes_alg for s, _, _ in recs) / rev_steps
    avg_step_s = rev_ms * 1e-3 / rev_steps
    achieved = step_bytes / avg_step_s / 1e9
    fwd_steps = sum(s.batches * P.nt for s, _, _ in recs)
    fwd_ms = sum(s.fwd_ms for s, _, _ in recs)
    fwd_bytes = sum(s.fwd_bytes_alg for s, _, _ in recs) / fwd_steps
    ratio = traffic_ratio("k_rev")
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if ratio is None else round(ratio * step_bytes),
            "traffic_source": None if ratio is None else
            f"HBM/algorithmic bytes {ratio} of k_rev from rocprofv3 FETCH_SIZE+WRITE_SIZE "
            "(profiles/k_fwd_k_rev_survey200_traffic.md) times this run's algorithmic bytes",
            "copy_gbs": None if copy_gbs is None else round(copy_gbs, 1),
            "kernel": "k_rev (reverse time step: reconstruction + receiver wavefield + imaging)",
            "avg_launch_us": round(avg_step_s * 1e6, 2), "alg_bytes_per_launch": step_bytes,
            "forward_step": {"kernel": "k_fwd", "achieved": round(fwd_bytes / (fwd_ms * 1e-3 / fwd_steps) / 1e9, 1),
                             "avg_launch_us": round(fwd_ms * 1e3 / fwd_steps, 2),
                             "alg_bytes_per_launch": fwd_bytes,
                             "note": "bytes of every tile of the step; tiles whose forward field is still +0 "
                                     "around them skip the step and move none of theirs, so this rate is the "
                                     "step's work per time, not an HBM rate"}}


def base_line(args, value, unit, elapsed, dtype, data, config, extra):
    out = {"metric": "grid-cell updates/s", "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(elapsed / max(args.steps, 1) * 1e3, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": data,
           "config": config}
    out.update(extra)
    return out


# ------------------------------------------------------------------ survey --
def run_survey(args, G, synth):
    P, vp, rho = synth.survey200(nshot=200)
    if args.nt:
        P.nt = args.nt
    S = args.shots_per_step
    store = {"f32": G.STORE_F32, "f16": G.STORE_F16, "bf16": G.STORE_BF16}[args.store]
    log(f"[bench] 200-shot survey: {P.nx}x{P.nz} nt={P.nt} npml={P.npml}, {S} shots/step, "
        f"steps={args.steps} warmup={args.warmup} store={args.store} chunk={args.chunk}")
    ctx = G.Context(0)
    copy_gbs = ctx.copy_probe(args.copy_mib << 20, 3) if args.copy_mib else None
    if copy_gbs:
        log(f"[bench] device-to-device copy of {args.copy_mib} MiB: {copy_gbs:.0f} GB/s (read + write)")
    mid = ctx.model_create(vp, rho)

    def step(k):
        first = 1 + (k * S) % 200
        P.startshot = first
        P.nshot = min(first + S - 1, 200)
        P.dshot = 1
        t0 = time.perf_counter()
        r = ctx.run(P, model=mid, batch=args.batch, store_fmt=store, graph_chunk=args.chunk, tile_z=args.tz,
                    want_post=True, ic_extra=args.ic_extra)
        return r, time.perf_counter() - t0, list(range(P.startshot, P.nshot + 1))

    for k in range(args.warmup):
        _, dt, _ = step(k)
        log(f"[bench] warmup step {k}: {dt:.2f} s")
    recs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r, dt, shots = step(args.warmup + k)
        recs.append((r.stats, dt, shots))
        s = r.stats
        log(f"[bench] step {k}: {len(shots)} shots {dt:.2f} s, {s.cell_updates / dt / 1e9:.1f} Gcell/s "
            f"(gpu {s.gpu_ms:.0f} ms, fwd {s.fwd_ms:.0f} rev {s.rev_ms:.0f}, tz {s.tile_z}/{s.tile_z_rev})")
    elapsed = time.perf_counter() - t0
    cells = sum(s.cell_updates for s, _, _ in recs)
    nshots = sum(len(sh) for _, _, sh in recs)
    roof = roofline_of(recs, P, copy_gbs)
    out = base_line(
        args, round(cells / elapsed / 1e9, 3), "Gcell/s", elapsed,
        "f32" if store == G.STORE_F32 else f"f32 stencil, {args.store} boundaries",
        "synthetic: seeded smooth-random layered model, data modelled in the same run",
        {"workload": "plain RTM, 200-shot survey, 1024x512 dx=10 m, PML 50, Ricker 15 Hz, dt=0.8 ms "
                     f"nt={P.nt}, shots every 50 m, split-spread 301 receivers, fp32, 3 imaging conditions"
                     + (" + the drop-in's decomposed, Poynting and energy-norm images" if args.ic_extra else ""),
         "shots_per_step": S, "batch": recs[0][0].batch, "graph_chunk": args.chunk,
         "tile_z": recs[0][0].tile_z, "tile_z_rev": recs[0][0].tile_z_rev, "store": args.store},
        {"shots_per_hour": round(nshots / elapsed * 3600.0, 1),
         "survey_latency_s": round(elapsed / max(args.steps, 1), 3),
         "hbm_pct_roofline": round(100.0 * roof["achieved"] / HBM_PEAK_GBS, 1),
         "roofline": roof, "cpu_baseline": None})
    gpu_shots = None
    if not args.quick:
        out["plain_form"] = plain_sub(args, G, ctx, P, mid, out["value"])
        out["single_shot"] = single_sub(args, G, synth, ctx)
        if not args.no_cpu:
            # the survey shots the CPU baseline runs, as the GPU images them
            Q = G.Params.from_buffer_copy(P)
            Q.startshot, Q.nshot, Q.dshot = CPU_SHOTS[0], CPU_SHOTS[-1], 1
            r = ctx.run(Q, model=mid, batch=len(CPU_SHOTS), graph_chunk=args.chunk, want_shots=True, want_post=False)
            gpu_shots = {CPU_SHOTS[k]: (r.shot_g1[k], r.shot_illum[k]) for k in range(len(CPU_SHOTS))}
    ctx.model_destroy(mid)
    if not args.quick:
        # the other configurations, as sub-objects of the same line
        out["batched"] = batched_sub(args, G, synth, ctx)
        out["precision"] = precision_sub(args, G, synth, ctx)
        out["served"] = served_sub(args, G, synth, ctx)
    ctx.close()
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle) ...")
        keep = {}
        out["cpu_baseline"] = cpu_baseline(P, vp, rho, synth, quick=args.quick, keep=keep)
        if gpu_shots:
            out["error_vs_oracle"] = survey_error(gpu_shots, keep)
        log(f"[bench] cpu baseline {out['cpu_baseline']}")
    return out


# ------------------------------------------------------ configuration subs --
def batched_sub(args, G, synth, ctx):
    """One 64-shot launch batch of the ragged 2048x768 survey (windows drawn
    from 400-1600 columns), graph chunk 50, autotuned tile heights: one
    untimed batch (it tunes), then one timed."""
    P, vp, rho, om = synth.batched_ragged()
    if args.nt:
        P.nt = args.nt
    P.startshot, P.nshot, P.dshot = 1, 64, 1
    mid = ctx.model_create(vp, rho)
    kw = dict(model=mid, batch=64, graph_chunk=args.chunk, tile_z=0, want_post=True, shot_offsetmax=om[:64])
    ctx.run(P, **kw)
    t0 = time.perf_counter()
    r = ctx.run(P, **kw)
    dt = time.perf_counter() - t0
    ctx.model_destroy(mid)
    roof = roofline_of([(r.stats, dt, None)], P)
    widths = [int(np.floor(om[i] / P.dx + 1)) for i in range(64)]
    out = {"workload": f"batched RTM, 2048x768 dx=10 m, shots 1-64 of 256 every 80 m, windows U[400,1600] "
                       f"columns (drawn {min(widths)}..{max(widths)}, clipped at the edges), nt={P.nt}, "
                       f"64 shots per launch batch, graph chunk {args.chunk}, autotuned tile heights",
           "value": round(r.stats.cell_updates / dt / 1e9, 3), "unit": "Gcell/s", "seconds": round(dt, 3),
           "shots_per_hour": round(64 / dt * 3600.0, 1), "tile_z": r.stats.tile_z, "tile_z_rev": r.stats.tile_z_rev,
           "k_rev_achieved_gbs": roof["achieved"], "k_rev_frac": roof["frac"]}
    log(f"[bench] batched: {out['value']} Gcell/s, k_rev {roof['frac']:.3f} of 8 TB/s")
    return out


def single_sub(args, G, synth, ctx):
    """The reference's CPU-runnable case on the GPU: one shot of the
    single-shot anomaly model (400x200, npml 40, nt 2000), as the plain form
    (one shot per launch, per-step launches, FP32 boundaries, 16-row tiles)
    and as the default engine (graph chunk, autotuned tile heights); one
    untimed run of each, then one timed."""
    P, vp, rho = synth.single_shot_anomaly()
    mid = ctx.model_create(vp, rho)
    out = {"workload": "single-shot anomaly model 400x200, dx=10 m, npml 40, nt=2000, 400 receivers, one shot, "
                       "images and post-processing included"}
    forms = (("plain_form", dict(batch=1, graph_chunk=0, tile_z=16, store_fmt=G.STORE_F32)),
             ("engine", dict(graph_chunk=args.chunk, tile_z=0)))
    for name, kw in forms:
        ctx.run(P, model=mid, want_post=True, **kw)
        t0 = time.perf_counter()
        r = ctx.run(P, model=mid, want_post=True, **kw)
        dt = time.perf_counter() - t0
        out[name] = {"s_per_shot": round(dt, 4), "gcell_s": round(r.stats.cell_updates / dt / 1e9, 3)}
        log(f"[bench] single shot, {name}: {dt * 1e3:.1f} ms, {out[name]['gcell_s']} Gcell/s")
    ctx.model_destroy(mid)
    return out


def run_single(args, G, synth):
    args.steps, args.warmup = 1, 1                      # one timed shot after one untimed
    ctx = G.Context(0)
    sub = single_sub(args, G, synth, ctx)
    ctx.close()
    e = sub["engine"]
    return base_line(args, e["gcell_s"], "Gcell/s", e["s_per_shot"], "f32",
                     "synthetic: the two-layer anomaly model, data modelled in the same run",
                     {"workload": sub["workload"], "shots_per_step": 1},
                     {"plain_form": sub["plain_form"], "cpu_baseline": None})


def plain_sub(args, G, ctx, P, mid, batched_value):
    """The plain one-shot-per-launch GPU form on survey shots 100-107: one shot
    per launch batch, per-step launches (no graph), FP32 boundaries, 16-row
    tiles."""
    Q = G.Params.from_buffer_copy(P)
    Q.startshot, Q.nshot, Q.dshot = 100, 107, 1
    n = Q.nshot - Q.startshot + 1
    ctx.run(Q, model=mid, batch=1, graph_chunk=0, tile_z=16, want_post=False)      # warm-up
    t0 = time.perf_counter()
    r = ctx.run(Q, model=mid, batch=1, graph_chunk=0, tile_z=16, want_post=True)
    dt = time.perf_counter() - t0
    v = r.stats.cell_updates / dt / 1e9
    log(f"[bench] plain form: {n} survey shots {dt:.2f} s, {v:.1f} Gcell/s")
    return {"shots": n, "gcell_s": round(v, 3), "shots_per_hour": round(n / dt * 3600.0, 1), "seconds": round(dt, 2),
            "note": "survey shots 100-107, one shot per launch batch, per-step launches, fp32 boundaries, 16-row tiles",
            "batched_vs_plain": round(batched_value / v, 2)}


def survey_error(gpu_shots, cpu_shots_done):
    """Relative L2 of the GPU's per-shot images against the FP32 oracle (the
    reference's operation order) on the survey shots the CPU baseline ran."""
    worst = {"g1": 0.0, "illum": 0.0}
    for ishot, (og1, oil) in cpu_shots_done.items():
        g1, il = gpu_shots[ishot]
        for k, a, b in (("g1", g1, og1), ("illum", il, oil)):
            e = float(np.linalg.norm((a - b).astype(np.float64)) / max(np.linalg.norm(b.astype(np.float64)), 1e-300))
            worst[k] = max(worst[k], e)
    out = {"sample": f"survey shots {sorted(cpu_shots_done)} at nt 4000, per-shot cross-correlation and illumination",
           "reference": "FP32 oracle in the reference's operation order", "max_rel_l2": {k: float(f"{v:.3e}") for k, v in worst.items()}}
    log(f"[bench] survey shots vs oracle: {out['max_rel_l2']}")
    return out


def precision_sub(args, G, synth, ctx):
    """2048x1024 dx=7.5 m 25 Hz dt=0.6 ms nt=6000 full-aperture shots per
    boundary store (FP32 / FP16 / BF16), and with FP16 / BF16 snapshots of the
    source wavefield every --snap-every steps on top of the 16-bit strips; one
    untimed batch of --prec-shots shots, then the same batch timed, per form
    (so that no slot is allocated in the timed run); image error of each form
    against the FP64 oracle on a shot the CPU affords (the same model,
    401-column window, nt 2000)."""
    P, vp, rho = synth.precision_grid()
    if args.nt:
        P.nt = args.nt
    S = args.prec_shots
    K = args.snap_every
    mid = ctx.model_create(vp, rho)
    # (name, boundary store, snapshot format or None)
    forms = [("f32", G.STORE_F32, None), ("f16", G.STORE_F16, None), ("bf16", G.STORE_BF16, None),
             ("f16+snap_f16", G.STORE_F16, G.STORE_F16), ("bf16+snap_bf16", G.STORE_BF16, G.STORE_BF16)]
    res = {}
    for name, fmt, snap in forms:
        kw = dict(store_fmt=fmt, graph_chunk=args.chunk)
        if snap is not None:
            kw.update(snap_every=K, snap_fmt=snap)
        Q = G.Params.from_buffer_copy(P)
        Q.startshot, Q.nshot, Q.dshot = 61, 60 + S, 1
        # a warm-up batch of the same shape: the timed run then reuses its
        # slots and graphs, as the survey's timed steps reuse the warm-up's
        ctx.run(Q, model=mid, batch=S, want_post=False, **kw)
        t0 = time.perf_counter()
        r = ctx.run(Q, model=mid, batch=S, want_post=True, **kw)
        dt = time.perf_counter() - t0
        res[name] = {"gcell_s": round(r.stats.cell_updates / dt / 1e9, 2), "shots_per_hour": round(S / dt * 3600, 1),
                     "seconds": round(dt, 2)}
        log(f"[bench] precision {name}: {S} shots {dt:.2f} s, {res[name]['gcell_s']} Gcell/s")
    snaps = {}
    for name in ("f16+snap_f16", "bf16+snap_bf16"):
        v = res.pop(name)
        v["vs_strips_alone"] = round(v["gcell_s"] / res[name.split("+")[0]]["gcell_s"], 4)
        snaps[name] = v
    out = {"workload": f"precision variant, 2048x1024 dx=7.5 m, PML 50, Ricker 25 Hz, dt=0.6 ms nt={P.nt}, "
                       f"full-aperture shots (2148x1124 padded cells), {S} shots per launch batch per store",
           "per_store": res,
           "snapshots": {"every": K, "note": "decimated source-wavefield snapshots (p, vx, vz of the interior, "
                                             "per-snapshot max-abs scale) restored in the reverse pass, on top of "
                                             "the boundary strips of the same precision", "per_form": snaps},
           "error_vs_fp64_oracle": None}
    if not args.no_cpu:
        from oracle import oracle as O

        Q = G.Params.from_buffer_copy(P)
        Q.startshot = Q.nshot = 64
        Q.flag_layout, Q.offsetmax, Q.nt = 1, 3000.0, min(P.nt, 2000)
        Qo = O.make_params(**{n: getattr(Q, n) for n, _ in Q._fields_})
        _omp_threads(_host_threads())
        t0 = time.perf_counter()
        ref = O.rtm(Qo, vp, rho, f64=True)
        refs = O.rtm(Qo, vp, rho, f64=True, snap_every=K)
        log(f"[bench] precision: FP64 oracle {time.perf_counter() - t0:.1f} s")

        def rel(x, y):
            return {k: float(f"{float(np.linalg.norm((a - b).astype(np.float64)) / max(np.linalg.norm(b.astype(np.float64)), 1e-300)):.3e}")
                    for k, a, b in zip(("image1", "image2s", "illum"), x, y)}

        err = {"sample": f"shot 64 of the precision model with a 401-column window (layout 1, offsetmax 3000 m), "
                         f"nt={Q.nt}",
               "note": "snapshot forms against the FP64 oracle with the same snapshots; the wavefield is still on "
                       "the grid at this nt, where the reference's reconstruction starts from zero source "
                       "velocities and departs from the forward field, which the snapshots restore: "
                       "snapshot_forms_vs_reference_semantics is that change of the images",
               "rel_l2": {}, "snapshot_forms_vs_reference_semantics": {}}
        for name, fmt, snap in forms:
            kw = {} if snap is None else dict(snap_every=K, snap_fmt=snap)
            r = ctx.run(Q, model=mid, batch=1, store_fmt=fmt, want_post=False, **kw)
            img = (r.image1, r.image2s, r.illum)
            err["rel_l2"][name] = rel(img, ref if snap is None else refs)
            log(f"[bench] precision {name} vs FP64 oracle: {err['rel_l2'][name]}")
            if snap is not None:
                err["snapshot_forms_vs_reference_semantics"][name] = rel(img, ref)
        out["error_vs_fp64_oracle"] = err
    ctx.model_destroy(mid)
    return out


def served_sub(args, G, synth, ctx):
    """A Poisson stream (4 req/s) of --requests requests over 8 resident
    models, served with continuous batching (FP16 boundaries), against the
    plain one-shot-per-launch GPU form (per-step launches, no graph, FP32
    boundaries, requests one after another) on the first --plain-requests."""
    models = synth.serve_models()
    reqs = synth.serve_stream(args.requests, models)
    mids = [ctx.model_create(v, r) for v, r in models]
    total_shots = sum(P.nshot for _, _, P in reqs)
    total_cells = 0.0
    for _, m, P in reqs:
        total_cells += cells_of(P, [G.geometry(P, s) for s in range(1, P.nshot + 1)])
    log(f"[bench] serve: {len(reqs)} requests, {total_shots} shots, {total_cells / 1e12:.2f} Tcell, "
        f"arrivals over {reqs[-1][0]:.1f} s")
    lat = []
    with G.Server(ctx, store_fmt=G.STORE_F16, mem_budget=args.mem_budget, tile_z=args.serve_tz) as srv:
        t0 = time.perf_counter()
        ids = []
        for (ta, m, P) in reqs:
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            ids.append((srv.submit(mids[m], P), P))
        for rid, P in ids:
            *_, st = srv.wait(rid, P, timeout=3600)
            lat.append(st.done_s - st.submit_s)
        served_s = time.perf_counter() - t0
        q = srv.query()
    log(f"[bench] served: {served_s:.1f} s, peak slots {q.peak_slots}, ticks {q.ticks}, "
        f"{q.allocations} device allocation(s) in {q.alloc_s:.2f} s (the budget reserved at server start)")
    nplain = min(args.plain_requests or len(reqs), len(reqs))
    plain_s, plain_shots, plain_cells, plain_lat = 0.0, 0, 0.0, []
    for (ta, m, P) in reqs[:nplain]:
        t1 = time.perf_counter()
        r = ctx.run(P, model=mids[m], batch=1, graph_chunk=0, tile_z=32, store_fmt=G.STORE_F32, want_post=True)
        d = time.perf_counter() - t1
        plain_s += d
        plain_shots += P.nshot
        plain_cells += r.stats.cell_updates
        plain_lat.append(d)
    log(f"[bench] plain: {nplain} requests in {plain_s:.1f} s")
    for m in mids:
        ctx.model_destroy(m)
    sph = total_shots / served_s * 3600.0
    psph = plain_shots / plain_s * 3600.0
    return {"workload": f"serving: Poisson stream 4 req/s, {len(reqs)} requests ({total_shots} shots, "
                        f"{total_cells / 1e12:.2f} Tcell), grids 256^2..2048x1024, nt U[1000,6000], 1-32 shots, "
                        "8 resident models, fp16 boundaries, continuous batching",
            "value": round(total_cells / served_s / 1e9, 3), "unit": "Gcell/s", "seconds": round(served_s, 2),
            "shots_per_hour": round(sph, 1), "requests": len(reqs),
            "p50_request_latency_s": round(float(np.percentile(lat, 50)), 2),
            "p99_request_latency_s": round(float(np.percentile(lat, 99)), 2), "peak_slots": q.peak_slots,
            "device_allocations": q.allocations, "allocation_s": round(q.alloc_s, 3),
            "plain_form": {"shots_per_hour": round(psph, 1), "gcell_s": round(plain_cells / plain_s / 1e9, 3),
                           "requests": nplain, "p50_request_s": round(float(np.percentile(plain_lat, 50)), 2),
                           "note": "one shot per launch, per-step launches, fp32 boundaries, requests one after "
                                   "another (the first requests of the same stream)"},
            "served_vs_plain_shots_per_hour": round(sph / psph, 2)}


# ------------------------------------------------------------------- field --
def run_field(args, G, synth):
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    P, vp, rho = synth.survey200(nshot=128)
    nsh, ntr, nt0, dt0, dx0 = 128, 151, 2000, 1.6e-3, 20.0
    P.mode, P.flag_layout, P.nshot, P.startshot, P.dshot = 1, 3, nsh, 1, 1
    if args.nt:
        P.nt = args.nt
        nt0 = P.nt // 2
    # one SEG-Y gather file: shot k at sx = 500 + 50 k m, traces at offsets 0, 20, ... 3000 m
    rng = np.random.default_rng(41)
    hdr = np.dtype([("tracl", "<i4"), ("p0", "V32"), ("offset", "<i4"), ("p1", "V32"), ("sx", "<i4"),
                    ("p2", "V38"), ("ns", "<i2"), ("dt", "<i2"), ("p3", "V122")])
    rec = np.zeros(nsh * ntr, dtype=[("h", hdr), ("d", "<f4", (nt0,))])
    rec["h"]["tracl"] = np.arange(1, nsh * ntr + 1)
    rec["h"]["offset"] = np.tile(np.arange(ntr) * int(dx0), nsh)
    rec["h"]["sx"] = np.repeat(500 + 50 * np.arange(nsh), ntr)
    rec["h"]["ns"], rec["h"]["dt"] = nt0, int(dt0 * 1e6)
    rec["d"] = rng.standard_normal((nsh * ntr, nt0)).astype(np.float32)
    tmpd = tempfile.mkdtemp(prefix="rtmh_field_")
    path = os.path.join(tmpd, "gather.sgy")
    with open(path, "wb") as f:
        f.write(b"\0" * 3600)
        rec.tofile(f)
    del rec
    log(f"[bench] field data: {nsh} gathers x {ntr} traces x {nt0} samples at {dx0:.0f} m / {dt0 * 1e3:.1f} ms "
        f"-> {P.nx}x{P.nz} dx={P.dx:.0f} nt={P.nt} dt={P.dt * 1e3:.1f} ms; steps={args.steps} warmup={args.warmup}")
    ctx = G.Context(0)
    mid = ctx.model_create(vp, rho)

    def step():
        # gathers are read and resampled on 16 host threads while the engine
        # runs the forward pass; it waits for shot k only when it needs k's record
        t0 = time.perf_counter()
        tab, endian = G.scan_gather(path, 1)
        fields, recs, jobs = [], [], []
        for k in range(tab.shape[1]):
            sx, n, itr, off = (int(v) for v in tab[:, k])
            offsetmax = np.float32(off + (n - 1) * dx0)
            f32 = np.float32
            n1, n2, no = int(f32(sx) / f32(P.dx)), int(f32(f32(sx) + offsetmax) / f32(P.dx)), int(f32(off) / f32(P.dx))
            n10, n20 = int(f32(sx) / f32(dx0)), int(f32(f32(sx) + offsetmax) / f32(dx0))
            fields.append((float(sx), float(off), float(offsetmax)))
            recs.append(np.zeros((P.nt, n2 - n1 + 1 - no), np.float32))
            jobs.append((itr, n, n20 - n10 + 1 - int(f32(off) / f32(dx0))))
        done = [0.0] * len(jobs)

        def one(k):
            itr, n, ncol0 = jobs[k]
            obs = G.read_gather(path, itr, n, nt0, ncol0, 1, endian)
            G.interp_seis(obs, recs[k].shape[1], P.nt, P.dt, dt0, out=recs[k])
            done[k] = time.perf_counter()

        ex = ThreadPoolExecutor(16)
        futs = [ex.submit(one, k) for k in range(len(jobs))]

        def wait(user, k):
            try:
                futs[k].result()
                return 0
            except Exception:
                return 1

        cb = G.SEIS_WAIT(wait)
        # rtm_real's batch rule: a quarter of the job, at least 32 shots
        batch = args.batch or (0 if nsh <= 32 else max(32, (nsh + 3) // 4))
        r = ctx.run(P, model=mid, seis_in=recs, shot_field=fields, batch=batch, graph_chunk=args.chunk,
                    tile_z=args.tz, want_post=True, seis_wait=cb)
        ex.shutdown()
        return r, time.perf_counter() - t0, max(done) - t0

    for k in range(args.warmup):
        _, dt, prep = step()
        log(f"[bench] warmup step {k}: {dt:.2f} s (host preparation done after {prep:.2f} s)")
    recs, preps = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r, dt, prep = step()
        recs.append((r.stats, dt, list(range(1, nsh + 1))))
        preps.append(prep)
        s = r.stats
        log(f"[bench] step {k}: {nsh} gathers {dt:.2f} s (host preparation done after {prep:.2f} s), "
            f"{s.cell_updates / dt / 1e9:.1f} Gcell/s whole, {s.cell_updates / (s.gpu_ms * 1e-3) / 1e9:.1f} on the GPU")
    elapsed = time.perf_counter() - t0
    cells = sum(s.cell_updates for s, _, _ in recs)
    gpu_s = sum(s.gpu_ms for s, _, _ in recs) * 1e-3
    roof = roofline_of(recs, P, None)
    out = base_line(
        args, round(cells / elapsed / 1e9, 3), "Gcell/s", elapsed, "f32",
        "synthetic: seeded random samples in a SEG-Y gather file written before the timed region",
        {"workload": f"field-data RTM (rtm_real path), survey model 1024x512 dx=10 m, PML 50, {nsh} one-sided "
                     f"gathers of {ntr} traces at 20 m and {nt0} samples at 1.6 ms resampled to 301 traces "
                     f"and nt={P.nt} at 0.8 ms, fp32",
         "shots_per_step": nsh, "batch": recs[0][0].batch, "graph_chunk": args.chunk, "tile_z": recs[0][0].tile_z,
         "tile_z_rev": recs[0][0].tile_z_rev},
        {"shots_per_hour": round(nsh * len(recs) / elapsed * 3600.0, 1),
         "host_preparation_done_after_s": round(float(np.mean(preps)), 3),
         "gpu_gcell_s": round(cells / gpu_s / 1e9, 3),
         "roofline": roof, "cpu_baseline": None})
    ctx.model_destroy(mid)
    ctx.close()
    os.remove(path)
    os.rmdir(tmpd)
    return out


# ----------------------------------------------------------------- batched --
def run_batched(args, G, synth):
    P, vp, rho, om = synth.batched_ragged()
    if args.nt:
        P.nt = args.nt
    S = args.shots_per_step or 64
    log(f"[bench] batched survey: {P.nx}x{P.nz} nt={P.nt}, {S} shots/step (ragged windows), steps={args.steps}")
    ctx = G.Context(0)
    mid = ctx.model_create(vp, rho)

    def step(k):
        first = 1 + (k * S) % 256
        last = min(first + S - 1, 256)
        Q = G.Params.from_buffer_copy(P)
        Q.startshot, Q.nshot, Q.dshot = first, last, 1
        t0 = time.perf_counter()
        r = ctx.run(Q, model=mid, batch=S, graph_chunk=args.chunk, tile_z=args.tz, want_post=True,
                    shot_offsetmax=om[first - 1:last])
        return r, time.perf_counter() - t0, list(range(first, last + 1))

    for k in range(args.warmup):
        _, dt, _ = step(k)
        log(f"[bench] warmup step {k}: {dt:.2f} s")
    recs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r, dt, shots = step(args.warmup + k)
        recs.append((r.stats, dt, shots))
        s = r.stats
        log(f"[bench] step {k}: {len(shots)} shots {dt:.2f} s, {s.cell_updates / dt / 1e9:.1f} Gcell/s "
            f"(gpu {s.gpu_ms:.0f} ms, fwd {s.fwd_ms:.0f} rev {s.rev_ms:.0f}, tz {s.tile_z}/{s.tile_z_rev})")
    elapsed = time.perf_counter() - t0
    cells = sum(s.cell_updates for s, _, _ in recs)
    nshots = sum(len(sh) for _, _, sh in recs)
    roof = roofline_of(recs, P)
    widths = [int(np.floor(om[i] / P.dx + 1)) for i in range(len(om))]
    out = base_line(
        args, round(cells / elapsed / 1e9, 3), "Gcell/s", elapsed, "f32",
        "synthetic: seeded smooth-random layered model, data modelled in the same run",
        {"workload": f"batched RTM, 2048x768 dx=10 m, 256 shots every 80 m, per-shot windows U[400,1600] columns "
                     f"(drawn {min(widths)}..{max(widths)}, clipped at the model edges), nt={P.nt}",
         "shots_per_step": S, "batch": S, "graph_chunk": args.chunk, "tile_z": recs[0][0].tile_z,
         "tile_z_rev": recs[0][0].tile_z_rev},
        {"shots_per_hour": round(nshots / elapsed * 3600.0, 1), "hbm_pct_roofline":
         round(100.0 * roof["achieved"] / HBM_PEAK_GBS, 1), "roofline": roof, "cpu_baseline": None})
    ctx.model_destroy(mid)
    ctx.close()
    return out


# --------------------------------------------------------------- precision --
def run_precision(args, G, synth):
    ctx = G.Context(0)
    sub = precision_sub(args, G, synth, ctx)
    ctx.close()
    best = sub["per_store"]["f16"]
    return base_line(
        args, best["gcell_s"], "Gcell/s", best["seconds"], "f32 stencil, f16 boundaries",
        "synthetic: seeded smooth-random layered model, data modelled in the same run",
        {"workload": sub["workload"], "shots_per_step": args.prec_shots, "batch": args.prec_shots},
        {"per_store": sub["per_store"], "snapshots": sub["snapshots"], "error_vs_fp64_oracle": sub["error_vs_fp64_oracle"],
         "cpu_baseline": None})


# ------------------------------------------------------------------- serve --
def run_serve(args, G, synth):
    ctx = G.Context(0)
    sub = served_sub(args, G, synth, ctx)
    ctx.close()
    extra = {k: v for k, v in sub.items() if k not in ("workload", "value", "unit", "seconds")}
    extra["cpu_baseline"] = None
    return base_line(args, sub["value"], "Gcell/s", sub["seconds"], "f32 stencil, f16 boundaries",
                     "synthetic: 8 seeded layered models, Poisson arrivals, data modelled in the same run",
                     {"workload": sub["workload"]}, extra)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (survey default 2, others 10)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="survey", choices=["survey", "batched", "precision", "serve", "field", "single"])
    ap.add_argument("--shots-per-step", type=int, default=0,
                    help="shots per step (survey default 200: the whole survey in one launch batch)")
    ap.add_argument("--batch", type=int, default=0, help="shots per launch (0: all shots of the step)")
    ap.add_argument("--store", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--chunk", type=int, default=50, help="time steps per replayed hipGraph (0: plain launches)")
    ap.add_argument("--tz", type=int, default=0, help="tile height 16/32/64 (0: autotuned)")
    ap.add_argument("--nt", type=int, default=0, help="override nt (shorter runs)")
    ap.add_argument("--requests", type=int, default=50, help="serve: requests in the stream")
    ap.add_argument("--plain-requests", type=int, default=8,
                    help="serve: first requests of the stream also run in the plain form (0: all)")
    ap.add_argument("--prec-shots", type=int, default=16, help="precision: timed shots per store")
    ap.add_argument("--snap-every", type=int, default=100, help="precision: steps between snapshots")
    ap.add_argument("--ic-extra", action="store_true",
                    help="survey: also the drop-in's other imaging conditions (decomposed, Poynting, energy norm)")
    ap.add_argument("--quick", action="store_true",
                    help="survey line only (no batched / precision / served sub-objects, no one-core baseline)")
    ap.add_argument("--mem-budget", type=float, default=0.0, help="serve: HBM budget in bytes (0: 80%% of free memory)")
    ap.add_argument("--serve-tz", type=int, default=0, help="serve: tile height 16/32/64 (0: the server's default)")
    ap.add_argument("--copy-mib", type=int, default=256,
                    help="survey: size of the device-to-device copy probe (achievable bandwidth; 0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / oracle error")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: every shot is independent and this build measures one GPU's "
              "throughput (the multi-GPU design shards shots and sums the FP64 stacks once; it is not run here).",
              file=sys.stderr)
        sys.exit(2)

    import gubingluo_amd as G
    from gubingluo_amd import synth

    if args.steps is None:
        args.steps = 2 if args.workload in ("survey", "field") else 10
    if args.workload == "survey":
        if not args.shots_per_step:
            args.shots_per_step = 200
        out = run_survey(args, G, synth)
    elif args.workload == "batched":
        out = run_batched(args, G, synth)
    elif args.workload == "precision":
        out = run_precision(args, G, synth)
    elif args.workload == "field":
        out = run_field(args, G, synth)
    elif args.workload == "single":
        out = run_single(args, G, synth)
    else:
        out = run_serve(args, G, synth)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
