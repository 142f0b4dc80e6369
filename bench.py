#!/usr/bin/env python3
"""bench.py -- 2-D acoustic prestack RTM on one MI355X.

Workload (the headline configuration): plain RTM of a 200-shot survey over a
seeded smooth-random layered model 1500-4500 m/s, nx=1024 nz=512, dx=10 m,
PML 50, Ricker 15 Hz, dt=0.8 ms nt=4000, shots every 50 m, split-spread 301
receivers (layout 1), fp32 stencil, all three imaging conditions
(cross-correlation, source-illumination normalised, Laplacian filtered).
Data: modelled by the same code from the synthetic model (no field data).

A "step" is the migration of one batch of --shots-per-step shots (forward
modelling with boundary saving, reverse reconstruction + receiver
propagation + imaging, FP64 stacking).  The velocity model is resident in
HBM before the timed region (rtmh_model_create); every step runs through the
C ABI (rtmh_run_model), which returns after its stream has drained.

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


def cpu_baseline(P_full, vp, rho, min_seconds=10.0, max_shots=6):
    """The oracle (FP32 restatement, OpenMP over rows) on a bounded sample of
    the same survey: whole shots (forward + reconstruction + receiver +
    imaging at the full nt) from the middle of the survey, until at least
    min_seconds of CPU work (at most max_shots)."""
    from oracle import oracle as O

    d = {n: getattr(P_full, n) for n, _ in P_full._fields_}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    cells = 0.0
    shots = []
    t0 = time.perf_counter()
    for k in range(max_shots):
        ishot = 100 + k
        d.update(nshot=ishot, startshot=ishot)
        Po = O.make_params(**d)
        g = O.geometry(Po, ishot)
        O.run_shot(Po, ishot, vp, rho)
        cells += 2.0 * Po.nt * g.nxp * g.nzp + (Po.nt - 1.0) * Po.nz * g.nxl
        shots.append(ishot)
        if time.perf_counter() - t0 >= min_seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(cells / dt / 1e9, 4), "unit": "Gcell/s", "cores": threads, "kind": "port",
            "sample": f"{len(shots)} whole shot(s) #{shots[0]}-#{shots[-1]} of the same survey (nt={P_full.nt}): "
                      f"{cells / 1e9:.2f} Gcell in {dt:.1f} s; {len(shots) / dt * 3600.0:.0f} shots/hour"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots-per-step", type=int, default=20)
    ap.add_argument("--batch", type=int, default=0, help="shots per launch (0: all shots of the step)")
    ap.add_argument("--store", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--chunk", type=int, default=50, help="time steps per replayed hipGraph (0: plain launches)")
    ap.add_argument("--tz", type=int, default=0, help="tile height (0: default)")
    ap.add_argument("--nt", type=int, default=0, help="override nt (shorter runs; the headline uses 4000)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: every shot is independent and this build measures one GPU's "
              "throughput (the multi-GPU design shards shots and sums the FP64 stacks once; it is not run here).",
              file=sys.stderr)
        sys.exit(2)

    import gubingluo_amd as G
    from gubingluo_amd import synth

    P, vp, rho = synth.cfg2(nshot=200)
    if args.nt:
        P.nt = args.nt
    S = args.shots_per_step
    store = {"f32": G.STORE_F32, "f16": G.STORE_F16, "bf16": G.STORE_BF16}[args.store]
    log(f"[bench] cfg2 survey: {P.nx}x{P.nz} nt={P.nt} npml={P.npml}, {S} shots/step, "
        f"steps={args.steps} warmup={args.warmup} store={args.store} chunk={args.chunk}")

    ctx = G.Context(0)
    mid = ctx.model_create(vp, rho)

    def step(k):
        first = 1 + (k * S) % 200
        P.startshot = first
        P.nshot = min(first + S - 1, 200)
        P.dshot = 1
        t0 = time.perf_counter()
        r = ctx.run(P, model=mid, batch=args.batch, store_fmt=store, graph_chunk=args.chunk, tile_z=args.tz,
                    want_post=True)
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
            f"(gpu {s.gpu_ms:.0f} ms, fwd {s.fwd_ms:.0f} rev {s.rev_ms:.0f})")
    elapsed = time.perf_counter() - t0

    cells = sum(s.cell_updates for s, _, _ in recs)
    nshots = sum(len(sh) for _, _, sh in recs)
    lat = np.concatenate([np.full(len(sh), dt) for _, dt, sh in recs])
    # dominant kernel: the reverse time step (k_vel_rev + k_pres_rev launched
    # back to back), timed by HIP events on the engine's stream over the
    # timed region; algorithmic bytes per step from the per-unit figures
    rev_steps = sum(s.batches * (P.nt - 1) for s, _, _ in recs)
    rev_ms = sum(s.rev_ms for s, _, _ in recs)
    step_bytes = float(np.mean([s.step_bytes_rev for s, _, _ in recs]))
    avg_step_s = rev_ms * 1e-3 / rev_steps
    achieved = step_bytes / avg_step_s / 1e9
    fwd_steps = sum(s.batches * P.nt for s, _, _ in recs)
    fwd_ms = sum(s.fwd_ms for s, _, _ in recs)
    fwd_bytes = float(np.mean([s.step_bytes_fwd for s, _, _ in recs]))

    out = {
        "metric": "grid-cell updates/s",
        "value": round(cells / elapsed / 1e9, 3),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if store == G.STORE_F32 else f"f32 stencil, {args.store} boundaries",
        "data": "synthetic: seeded smooth-random layered model, data modelled in the same run",
        "config": {"workload": "plain RTM, 200-shot survey, 1024x512 dx=10 m, PML 50, Ricker 15 Hz, dt=0.8 ms "
                               f"nt={P.nt}, shots every 50 m, split-spread 301 receivers, fp32, 3 imaging conditions",
                   "shots_per_step": S, "batch": recs[0][0].batch, "graph_chunk": args.chunk,
                   "tile_z": recs[0][0].tile_z, "store": args.store},
        "shots_per_hour": round(nshots / elapsed * 3600.0, 1),
        "p50_shot_latency_s": round(float(np.percentile(lat, 50)), 3),
        "p99_shot_latency_s": round(float(np.percentile(lat, 99)), 3),
        "hbm_pct_roofline": round(100.0 * achieved / HBM_PEAK_GBS, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "reverse time step (k_vel_rev + k_pres_rev)",
                     "avg_launch_us": round(avg_step_s * 1e6, 2), "alg_bytes_per_launch": step_bytes,
                     "forward_step": {"achieved": round(fwd_bytes / (fwd_ms * 1e-3 / fwd_steps) / 1e9, 1),
                                      "avg_launch_us": round(fwd_ms * 1e3 / fwd_steps, 2),
                                      "alg_bytes_per_launch": fwd_bytes}},
        "cpu_baseline": None,
    }
    ctx.model_destroy(mid)
    ctx.close()
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(P, vp, rho)
        log(f"[bench] cpu baseline {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
