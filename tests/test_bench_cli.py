"""bench.py on the CPU: the one-GPU contract (more GPUs exit with a message)
and the synthetic workload generators it measures (shapes, seeds, windows)."""
import os
import subprocess
import sys

import numpy as np

import gubingluo_amd as G
from gubingluo_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_survey200_geometry():
    P, vp, rho = synth.survey200(nshot=200)
    assert vp.shape == (512, 1024) and vp.min() >= 1500.0 and vp.max() <= 4725.0
    g = G.geometry(P, 100)
    assert g.nrec == 301 and g.nxl == 301 and g.nzp == 612       # split spread, layout 1


def test_batched_ragged_windows():
    P, vp, rho, om = synth.batched_ragged()
    assert vp.shape == (768, 2048) and len(om) == 256
    NX = np.floor(om / P.dx + 1).astype(int)
    widths = 2 * ((NX - 1) // 2) + 1                               # layout-1 windows are odd
    assert NX.min() >= 400 and NX.max() <= 1600
    # seeded: the same draw every time
    assert np.array_equal(om, synth.batched_ragged()[3])
    # a window per shot through the same geometry rule
    Pk = G.Params.from_buffer_copy(P)
    Pk.offsetmax = float(om[127])
    g = G.geometry(Pk, 128)
    assert g.nxl == int(widths[127])


def test_precision_grid_and_serve_shapes():
    P, vp, rho = synth.precision_grid()
    assert vp.shape == (1024, 2048) and P.nt == 6000 and abs(P.dt - 6e-4) < 1e-9 and P.dx == 7.5
    assert G.geometry(P, 1).nxl == 2048
    models = synth.serve_models()
    assert len(models) == 8
    for vpm, _ in models:
        assert 256 <= vpm.shape[0] <= 1024 and 256 <= vpm.shape[1] <= 2048
    reqs = synth.serve_stream(50, models)
    t = np.array([r[0] for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert all(1000 <= r[2].nt <= 6000 and 1 <= r[2].nshot <= 32 for r in reqs)


def test_survey_error_is_the_worst_shot():
    sys.path.insert(0, ROOT)
    import bench

    rng = np.random.default_rng(3)
    ref = {k: (rng.standard_normal((8, 5)).astype(np.float32), rng.random((8, 5)).astype(np.float32)) for k in (1, 2)}
    gpu = {k: (a.copy(), b.copy()) for k, (a, b) in ref.items()}
    gpu[2][0][0, 0] += 0.5                               # one shot off in g1 only
    e = bench.survey_error(gpu, ref)["max_rel_l2"]
    want = np.linalg.norm(gpu[2][0] - ref[2][0]) / np.linalg.norm(ref[2][0])
    assert abs(e["g1"] - want) <= 1e-3 * want and e["illum"] == 0.0


def test_every_workload_is_a_choice():
    """each documented workload parses (the one-GPU check runs before any GPU
    work, so --gpus 2 ends every one of them with the same message)"""
    for w in ("survey", "batched", "precision", "serve", "field", "single"):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--workload", w, "--gpus", "2"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 2 and "one GPU" in r.stderr, (w, r.stderr[-300:])
