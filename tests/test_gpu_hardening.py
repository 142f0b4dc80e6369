"""Guards of the engine and the served form on the GPU, through the C ABI:

  graph cache     batches with equal tile counts and different slot counts
                  replay their own graphs (the slot count is part of the key):
                  bit for bit equal to plain launches
  stability       dt above min(dx,dz)/(sqrt(2) sum|c| vmax) is refused
                  (get_constant, RTM_GPU_velstr.cpp:355-361)
  blow-up         a wavefield that stops being finite ends the job (engine)
                  or the request (server) with status -7 (the reference's
                  max|p| print every ntsnap steps, rtm_model.cu:271-273)
  budget          a request whose shot cannot fit the server's budget is
                  refused at submit; one that fits one shot at a time runs
                  one shot at a time; the budget is reserved in one device
                  allocation at start (on demand when the device cannot
                  give it in one piece)
  field batches   recorded gathers migrated in unequal batches with graph
                  replay and the record callback equal one batch of plain
                  launches and the kernel-order oracle, gather by gather
"""
import numpy as np
import pytest

import gubingluo_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def model(nz, nx, seed, hot=None):
    rng = np.random.default_rng(seed)
    vp = np.full((nz, nx), 2000.0, np.float32)
    vp[nz // 2:] = 2600.0
    vp += rng.uniform(-50, 50, (nz, nx)).astype(np.float32)
    rho = (2000.0 + rng.uniform(-100, 100, (nz, nx))).astype(np.float32)
    if hot is not None:
        rho[hot] = 1e37        # finite, but rho v^2 overflows to infinity
    return vp, rho


def kw_base(**kw):
    d = dict(nshot=1, startshot=1, dshot=1, flag_layout=2, nx=120, nz=60, nt=300, dx=10.0, dz=10.0, dt=1e-3,
             npml=32, fdom=20.0, spx0=600.0, spz0=20.0, dspx=50.0, offsetmax=1200.0, tsnap=0.05)
    d.update(kw)
    return d


@pytest.mark.parametrize("fmt", [G.STORE_F32, G.STORE_F16])
def test_graph_cache_keyed_by_slot_count(gpu_ctx, fmt):
    """Shot 1 has a 225-column padded window (4 tile columns), shots 2 and 3
    105-column ones (2 each); a memory budget between 2 small slots and one
    big plus one small slot batches them as [1] and [2, 3]: equal tile counts,
    1 slot and then 2.  A graph captured for the first batch would save and
    advance only one slot of the second."""
    nz, nx = 60, 400
    vp, rho = model(nz, nx, 5)
    kw = kw_base(nshot=3, flag_layout=1, nx=nx, nz=nz, nt=300, spx0=2000.0, dspx=500.0)
    P = G.make_params(**kw)
    om = np.array([1600.0, 400.0, 400.0], np.float32)
    g = [G.geometry(G.make_params(**dict(kw, offsetmax=float(o))), k + 1) for k, o in enumerate(om)]
    cols = [(x.nxp + 63) // 64 for x in g]
    assert cols[0] == cols[1] + cols[2], cols
    mid = gpu_ctx.model_create(vp, rho)
    try:
        with G.Server(gpu_ctx, max_slots=4, store_fmt=fmt) as probe:   # slot sizes (+ a few KB of images)
            size = []
            for k, o in enumerate(om):
                Q = G.make_params(**dict(kw, offsetmax=float(o), startshot=k + 1, nshot=k + 1))
                size.append(probe.request_bytes(mid, Q)[0])
        big, small = size[0], size[1]
        budget = 0.5 * (2 * small + big + small)
        assert 2 * small < budget < big + small
        print(f"[graph key] store {fmt}, tile columns {cols}, slots {big / 1e6:.1f}/{small / 1e6:.1f} MB", flush=True)
        ref = gpu_ctx.run(P, model=mid, graph_chunk=0, tile_z=16, store_fmt=fmt, want_shots=True,
                          shot_offsetmax=om, mem_budget=budget)
        assert ref.stats.batches == 2
        for rep in range(2):                   # the second run replays both cached graphs
            r = gpu_ctx.run(P, model=mid, graph_chunk=20, tile_z=16, store_fmt=fmt, want_shots=True,
                            shot_offsetmax=om, mem_budget=budget)
            assert r.stats.batches == 2
            for k in range(3):
                assert np.array_equal(r.shot_seis[k], ref.shot_seis[k]), (rep, k)
                assert np.array_equal(r.shot_g1[k], ref.shot_g1[k]), (rep, k)
                assert np.array_equal(r.shot_illum[k], ref.shot_illum[k]), (rep, k)
            assert np.array_equal(r.image1, ref.image1) and np.array_equal(r.image2s, ref.image2s)
    finally:
        gpu_ctx.model_destroy(mid)
    if fmt == G.STORE_F32:
        Po = O.make_params(**kw)
        for k in range(3):
            Po.offsetmax = float(om[k])
            o = O.run_shot(Po, k + 1, vp, rho, korder=True)
            assert np.array_equal(ref.shot_g1[k], o.g1), k


def _dt_stable(P, vmax):
    c = O.coef_table()[90:100].astype(np.float32)
    s = np.float32(0.0)
    for x in c:
        s = np.float32(s + np.abs(x))
    h = np.float32(min(P.dx, P.dz))
    return float(h / (np.float32(np.sqrt(np.float32(2.0))) * s * np.float32(vmax)))


def test_stability_bound_refused(gpu_ctx):
    vp, rho = model(60, 120, 6)
    P = G.make_params(**kw_base())
    dtm = _dt_stable(P, float(vp.max()))
    print(f"[stability] bound {dtm:.4e} s", flush=True)
    P.dt = dtm * 1.05
    with pytest.raises(RuntimeError, match="stability bound"):
        gpu_ctx.run(P, vp, rho)
    P.dt = dtm * 0.95
    r = gpu_ctx.run(P, vp, rho)
    assert np.isfinite(r.image1).all()


def test_blow_up_ends_the_job(gpu_ctx):
    """rho v^2 overflows in a block of cells: the forward wavefield stops
    being finite and the job ends with -7 naming the shot."""
    vp, rho = model(60, 120, 7, hot=(slice(40, 44), slice(50, 54)))
    P = G.make_params(**kw_base(nshot=2, dshot=1, dspx=100.0))
    print("[blow-up] engine", flush=True)
    for chunk in (0, 25):
        with pytest.raises(RuntimeError, match=r"\(-7\).*not finite"):
            gpu_ctx.run(P, vp, rho, graph_chunk=chunk)
    # the engine stays usable
    vp2, rho2 = model(60, 120, 7)
    r = gpu_ctx.run(P, vp2, rho2, graph_chunk=25)
    assert np.isfinite(r.image1).all()


def test_non_finite_model_refused(gpu_ctx):
    vp, rho = model(60, 120, 8)
    vp[3, 4] = np.nan
    with pytest.raises(RuntimeError, match="bad model"):
        gpu_ctx.model_create(vp, rho)
    vp, rho = model(60, 120, 8)
    rho[5, 5] = 0.0
    with pytest.raises(RuntimeError, match="bad model"):
        gpu_ctx.model_create(vp, rho)


def test_server_budget_and_blow_up(gpu_ctx):
    vp, rho = model(60, 160, 9)
    vb, rb = model(60, 160, 9, hot=(slice(40, 44), slice(70, 74)))
    kw = kw_base(nshot=3, flag_layout=1, nx=160, spx0=300.0, dspx=300.0, offsetmax=800.0)
    P = G.make_params(**kw)
    mid = gpu_ctx.model_create(vp, rho)
    mbad = gpu_ctx.model_create(vb, rb)
    try:
        ref = gpu_ctx.run(P, model=mid, batch=3)
        with G.Server(gpu_ctx, max_slots=4) as probe:
            shot, req = probe.request_bytes(mid, P)
        print(f"[server budget] shot {shot / 1e6:.1f} MB, request block {req / 1e6:.1f} MB", flush=True)
        # too small for one shot: refused at submit, nothing queued
        with G.Server(gpu_ctx, max_slots=4, mem_budget=req + 0.5 * shot) as srv:
            with pytest.raises(RuntimeError, match="budget"):
                srv.submit(mid, P)
            assert srv.query().queued_requests == 0
        # one shot at a time
        with G.Server(gpu_ctx, max_slots=4, mem_budget=req + 1.5 * shot) as srv:
            rid = srv.submit(mid, P)
            i1, i2, il, lap, nrm, st = srv.wait(rid, P, timeout=120)
            assert st.nshots == 3 and np.array_equal(i1, ref.image1) and np.array_equal(il, ref.illum)
            assert srv.query().peak_slots == 1
        # a request that blows up fails alone; the one beside it completes
        with G.Server(gpu_ctx, max_slots=8) as srv:
            bad = srv.submit(mbad, P)
            good = srv.submit(mid, P)
            with pytest.raises(RuntimeError, match=r"\(-7\).*not finite"):
                srv.wait(bad, P, timeout=120)
            i1, *_rest, st = srv.wait(good, P, timeout=120)
            assert np.array_equal(i1, ref.image1)
        # dt above the stability bound: refused at submit
        with G.Server(gpu_ctx, max_slots=4) as srv:
            Q = G.make_params(**dict(kw, dt=2.0 * _dt_stable(P, float(vp.max()))))
            with pytest.raises(RuntimeError, match="stability bound"):
                srv.submit(mid, Q)
    finally:
        gpu_ctx.model_destroy(mid)
        gpu_ctx.model_destroy(mbad)


def test_server_arena_and_on_demand_fallback(gpu_ctx):
    """A server takes its budget in one device allocation at start and serves
    every request block and slot from it (one allocation in all); a budget
    the device cannot give in one piece (here 4 TB) leaves the server
    allocating as shots are admitted, with the same images."""
    vp, rho = model(60, 160, 11)
    kw = kw_base(nshot=3, flag_layout=1, nx=160, spx0=300.0, dspx=300.0, offsetmax=800.0)
    P = G.make_params(**kw)
    mid = gpu_ctx.model_create(vp, rho)
    try:
        ref = gpu_ctx.run(P, model=mid, batch=3)
        print("[server arena] default budget", flush=True)
        with G.Server(gpu_ctx, max_slots=4) as srv:
            ids = [srv.submit(mid, P) for _ in range(2)]
            for rid in ids:
                i1, i2, il, *_rest = srv.wait(rid, P, timeout=120)
                assert np.array_equal(i1, ref.image1) and np.array_equal(il, ref.illum)
            q = srv.query()
            assert q.allocations == 1, q.allocations
            assert q.budget_bytes > 0 and int(q.budget_bytes) % 256 == 0
        print("[server arena] 4 TB budget: on demand", flush=True)
        with G.Server(gpu_ctx, max_slots=4, mem_budget=4e12) as srv:
            q0 = srv.query()
            assert q0.budget_bytes == 4e12 and q0.allocations == 1     # the reservation that failed
            rid = srv.submit(mid, P)
            i1, i2, il, *_rest = srv.wait(rid, P, timeout=120)
            assert np.array_equal(i1, ref.image1) and np.array_equal(il, ref.illum)
            assert srv.query().allocations > 1
        # the refused reservation left no error behind for the next job
        r = gpu_ctx.run(P, model=mid, batch=3)
        assert np.array_equal(r.image1, ref.image1)
    finally:
        gpu_ctx.model_destroy(mid)


def test_server_budget_counts_the_context_cached_memory(gpu_ctx):
    """Slot memory that the context's jobs keep cached for reuse goes back to
    the device when a server starts, so the default budget (80 % of the free
    memory) is the same right after a large job as on an idle context."""
    from gubingluo_amd import synth

    P, vp, rho = synth.survey200(nshot=200)
    P.startshot, P.nshot, P.dshot = 1, 24, 1                  # 24 survey slots at nt 4000: ~13 GB
    print("[server budget] 24 survey shots, then two servers", flush=True)
    mid = gpu_ctx.model_create(vp, rho)
    try:
        gpu_ctx.run(P, model=mid, want_post=False)
        with G.Server(gpu_ctx, max_slots=4) as a:
            b1 = a.query().budget_bytes
        with G.Server(gpu_ctx, max_slots=4) as b:
            b2 = b.query().budget_bytes
    finally:
        gpu_ctx.model_destroy(mid)
    print(f"[server budget] after the job {b1 / 1e9:.2f} GB, idle {b2 / 1e9:.2f} GB", flush=True)
    assert b1 >= 0.995 * b2


def test_job_after_a_failed_pool_allocation():
    """A job whose slots do not fit beside the blocks an earlier job left
    cached: the pool's first hipMalloc fails, the cached blocks go back to the
    device and the second try succeeds -- and that handled failure must not
    surface as the error of a later kernel launch.  Survey shots with source
    snapshots at every step hold ~7.4 GB (FP32) or ~4 GB (FP16) per slot, so
    24 FP32 slots (~177 GB cached, not reusable for the smaller FP16 slots)
    leave too little for 32 FP16 ones (~128 GB)."""
    from gubingluo_amd import synth

    P, vp, rho = synth.survey200(nshot=200)
    ctx = G.Context(0)                                   # its own pool, released at the end
    try:
        mid = ctx.model_create(vp, rho)
        P.startshot, P.nshot, P.dshot = 1, 24, 1
        print("[pool] 24 shots, FP32 snapshots every step", flush=True)
        ctx.run(P, model=mid, snap_every=1, snap_fmt=G.STORE_F32, graph_chunk=50, want_post=False)
        P.nshot = 32
        print("[pool] 32 shots, FP16 snapshots every step", flush=True)
        r = ctx.run(P, model=mid, snap_every=1, snap_fmt=G.STORE_F16, graph_chunk=50, want_post=False)
        assert np.isfinite(r.image1).all() and np.abs(r.image1).max() > 0
        ctx.model_destroy(mid)
    finally:
        ctx.close()


def test_field_gathers_in_unequal_batches(gpu_ctx):
    """rtm_real's batching (batches of the job, the record callback before
    each gather's reverse pass, graph replay): five gathers in batches of
    2, 2, 1 with chunk 16 equal one batch of plain launches and the
    kernel-order oracle, gather by gather."""
    nz, nx, nt = 50, 220, 240
    vp, rho = model(nz, nx, 31)
    fields = [(123.0, 0.0, 600.0), (487.5, 35.0, 905.0), (1000.0, 20.0, 700.0), (640.0, 10.0, 500.0),
              (300.0, 0.0, 1100.0)]
    kw = dict(nshot=5, startshot=1, dshot=1, mode=1, flag_layout=3, nx=nx, nz=nz, nt=nt, dx=10.0, dz=10.0,
              dt=1e-3, npml=32, fdom=20.0, spz0=20.0, direct=1.5, direct0=0.05)
    P = G.make_params(**kw)
    rng = np.random.default_rng(32)
    geoms = [G.gather_geometry(P, k + 1, f) for k, f in enumerate(fields)]
    recs = [rng.standard_normal((nt, g.nrec)).astype(np.float32) for g in geoms]
    calls = []
    cb = G.SEIS_WAIT(lambda user, k: calls.append(k) or 0)
    print("[field batches] 5 gathers, batch 2, chunk 16", flush=True)
    a = gpu_ctx.run(P, vp, rho, want_shots=True, seis_in=recs, shot_field=fields, seis_wait=cb, batch=2,
                    graph_chunk=16)
    assert a.stats.batches == 3 and calls == [0, 1, 2, 3, 4]
    b = gpu_ctx.run(P, vp, rho, want_shots=True, seis_in=recs, shot_field=fields, batch=0, graph_chunk=0)
    assert b.stats.batches == 1
    for k, f in enumerate(fields):
        assert np.array_equal(a.shot_g1[k], b.shot_g1[k]), k
        assert np.array_equal(a.shot_illum[k], b.shot_illum[k]), k
        Po = O.make_params(**dict(kw, field=1, fsx=f[0], foffset=f[1], foffsetmax=f[2]))
        o = O.run_shot(Po, k + 1, vp, rho, seis_in=recs[k], korder=True)
        assert np.array_equal(a.shot_g1[k], o.g1), k
        assert np.array_equal(a.shot_illum[k], o.illum), k
    assert np.array_equal(a.image1, b.image1) and np.array_equal(a.image2s, b.image2s)
