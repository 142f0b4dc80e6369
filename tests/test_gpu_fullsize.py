"""GPU parity at the full size of the 200-shot survey (1024x512 model, dx 10 m,
PML 50, windows of up to 401 x 612 padded cells, nt 4000): whole shots of the
survey against the CPU oracle (bit-exact against the kernel-order form, the
seismogram and image bars against the reference order), and the properties
that hold at any size: a shot's result does not depend on the batch it runs
in or on the tile height, and the 16-bit boundary stores stay within the
reduced-precision bar of the fp32 store."""
import numpy as np
import pytest

import gubingluo_amd as G
from gubingluo_amd import synth
from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def survey():
    P, vp, rho = synth.survey200(nshot=200)
    return {n: getattr(P, n) for n, _ in P._fields_}, vp, rho


def shots_of(d, first, last):
    d = dict(d, startshot=first, nshot=last)
    return G.make_params(**d), O.make_params(**d)


@pytest.mark.parametrize("ishot", [1, 100, 200])
def test_survey_shot_full_size_vs_oracle(gpu_ctx, survey, ishot):
    """shot 1 and 200: windows clipped at the model edges; shot 100: full aperture"""
    d, vp, rho = survey
    Pg, Po = shots_of(d, ishot, ishot)
    g = G.geometry(Pg, ishot)
    print(f"[fullsize] shot {ishot} window {g.nxp}x{g.nzp} nt {Pg.nt}", flush=True)
    r = gpu_ctx.run(Pg, vp, rho, want_shots=True)
    k = O.run_shot(Po, ishot, vp, rho, korder=True)
    assert np.array_equal(r.shot_seis[0], k.seis)
    assert np.array_equal(r.shot_g1[0], k.g1)
    assert np.array_equal(r.shot_illum[0], k.illum)
    if ishot == 100:
        o = O.run_shot(Po, ishot, vp, rho)
        assert rel_l2(r.shot_seis[0], o.seis) <= 1e-5
        assert rel_l2(r.shot_g1[0], o.g1) <= 1e-4
        assert rel_l2(r.shot_illum[0], o.illum) <= 1e-4


def test_survey_batch_and_tile_invariance_full_size(gpu_ctx, survey):
    """shots 1-6 (growing windows) alone, 16-row tiles, against one batch of six
    with 64-row tiles: every shot's seismogram and images identical"""
    d, vp, rho = survey
    Pg, _ = shots_of(d, 1, 6)
    print("[fullsize] shots 1-6: batch 1 tile 16 vs batch 6 tile 64", flush=True)
    a = gpu_ctx.run(Pg, vp, rho, batch=1, tile_z=16, want_shots=True)
    b = gpu_ctx.run(Pg, vp, rho, batch=6, tile_z=64, graph_chunk=50, want_shots=True)
    for k in range(6):
        assert np.array_equal(a.shot_seis[k], b.shot_seis[k]), k
        assert np.array_equal(a.shot_g1[k], b.shot_g1[k]), k
        assert np.array_equal(a.shot_illum[k], b.shot_illum[k]), k
    assert np.array_equal(a.image1, b.image1) and np.array_equal(a.image2s, b.image2s)


@pytest.mark.parametrize("fmt", [G.STORE_F16, G.STORE_BF16])
def test_survey_16bit_store_full_size(gpu_ctx, survey, fmt):
    """shot 100 over nt 4000 with 16-bit boundary stores against the fp32 store:
    the seismogram is untouched (forward pass), the images within 1e-3"""
    d, vp, rho = survey
    Pg, _ = shots_of(d, 100, 100)
    print(f"[fullsize] shot 100 store {fmt}", flush=True)
    a = gpu_ctx.run(Pg, vp, rho, want_shots=True)
    b = gpu_ctx.run(Pg, vp, rho, want_shots=True, store_fmt=fmt)
    assert np.array_equal(a.shot_seis[0], b.shot_seis[0])
    assert rel_l2(b.shot_g1[0], a.shot_g1[0]) <= 1e-3
    assert rel_l2(b.shot_illum[0], a.shot_illum[0]) <= 1e-3
    assert rel_l2(b.image1, a.image1) <= 1e-3
