"""Served form (continuous batching): each request's images equal the batch
engine's result for the same job, bit for bit, whether the request runs
alone or mixed with requests of other grid sizes, nt, models and arrival
times (shots at different phases and time steps share launches)."""
import time

import numpy as np
import pytest

import gubingluo_amd as G

pytestmark = pytest.mark.gpu


def model(nz, nx, seed):
    rng = np.random.default_rng(seed)
    vp = np.full((nz, nx), 1900.0, np.float32)
    vp[nz // 2:] = 2600.0
    vp += rng.uniform(-80, 80, (nz, nx)).astype(np.float32)
    return vp, np.full((nz, nx), 2000.0, np.float32)


JOBS = [
    dict(nx=160, nz=60, nt=300, nshot=3, dshot=1, flag_layout=1, spx0=300.0, dspx=300.0, offsetmax=800.0, npml=32),
    dict(nx=120, nz=80, nt=420, nshot=2, dshot=1, flag_layout=2, spx0=400.0, dspx=300.0, npml=40, fdom=15.0),
    dict(nx=200, nz=50, nt=260, nshot=5, dshot=2, flag_layout=3, spx0=100.0, dspx=200.0, offset=50.0,
         offsetmax=600.0, npml=36),
]


@pytest.fixture(scope="module")
def setup(gpu_ctx):
    mods, params, alone = [], [], []
    for k, kw in enumerate(JOBS):
        vp, rho = model(kw["nz"], kw["nx"], k)
        mods.append(gpu_ctx.model_create(vp, rho))
        P = G.make_params(**dict(dict(dx=10.0, dz=10.0, dt=1e-3, fdom=20.0, spz0=20.0, startshot=1), **kw))
        params.append(P)
    yield mods, params
    for m in mods:
        gpu_ctx.model_destroy(m)


@pytest.mark.parametrize("fmt", [G.STORE_F32, G.STORE_F16])
def test_served_alone_and_mixed_equal_batch_engine(gpu_ctx, setup, fmt):
    mods, params = setup
    print(f"[server] store_fmt={fmt}", flush=True)
    ref = [gpu_ctx.run(P, model=m, store_fmt=fmt, batch=2) for m, P in zip(mods, params)]
    with G.Server(gpu_ctx, store_fmt=fmt, max_slots=6) as srv:
        # alone, one at a time
        for m, P, r in zip(mods, params, ref):
            rid = srv.submit(m, P)
            i1, i2, il, lap, nrm, st = srv.wait(rid, P, timeout=120)
            assert np.array_equal(i1, r.image1) and np.array_equal(i2, r.image2s) and np.array_equal(il, r.illum)
            assert np.array_equal(lap, r.image_lap) and np.array_equal(nrm, r.image_norm)
        # mixed: staggered arrivals so shots of different requests overlap in every phase
        ids = []
        for m, P in zip(mods, params):
            ids.append(srv.submit(m, P))
            time.sleep(0.01)
        ids.append(srv.submit(mods[0], params[0]))
        for rid, (m, P), r in zip(ids, list(zip(mods, params)) + [(mods[0], params[0])], ref + [ref[0]]):
            i1, i2, il, lap, nrm, st = srv.wait(rid, P, timeout=120)
            assert np.array_equal(i1, r.image1) and np.array_equal(i2, r.image2s) and np.array_equal(il, r.illum)
        q = srv.query()
        print(f"[server] ticks={q.ticks} peak_slots={q.peak_slots}", flush=True)
        assert q.peak_slots >= 3


@pytest.mark.parametrize("fmt", [G.STORE_F32, G.STORE_F16])
def test_served_vs_oracle(gpu_ctx, setup, fmt):
    """Served requests against the kernel-order oracle's job with the same
    boundary store (FP32, and FP16 as the served benchmark runs), bit for bit;
    the FP32 store also against the reference order within the image bar."""
    from oracle import oracle as O
    from tests.conftest import rel_l2

    mods, params = setup
    with G.Server(gpu_ctx, max_slots=6, store_fmt=fmt) as srv:
        ids = [srv.submit(m, P) for m, P in zip(mods, params)]
        for k, (rid, P) in enumerate(zip(ids, params)):
            i1, i2, il, lap, nrm, st = srv.wait(rid, P, timeout=120)
            vp, rho = model(P.nz, P.nx, k)
            Po = O.make_params(**{n: getattr(P, n) for n, _ in P._fields_})
            k1, k2, kl = O.rtm(Po, vp, rho, korder=True, store_fmt=fmt)
            assert np.array_equal(i1, k1) and np.array_equal(i2, k2) and np.array_equal(il, kl), k
            if fmt == G.STORE_F32:
                f1, f2, fl = O.rtm(Po, vp, rho)
                assert rel_l2(i1, f1) <= 1e-4 and rel_l2(i2, f2) <= 1e-4 and rel_l2(il, fl) <= 1e-4, k
            print(f"[server] store {fmt} request {k} bit-exact vs the kernel-order oracle", flush=True)
