"""Host logic of librtm_hip.so that needs no GPU, checked on the CPU by a
small C++ program (tests/host_logic.cpp) linked against the library: the
served form's admission rules driven by a fake device allocator (a shot
larger than the budget is refused, never queued forever; a device that
refuses memory with nothing resident fails the request; one-shot budgets
admit in arrival order), the server's first-fit device arena (blocks never
overlap, the lowest fitting hole is taken, freed neighbours merge), the
per-XCD runs of the tile table (every tile once, long tiles first, padding at
run ends, runs of equal cost, each a contiguous piece, the padded table within
the engine's capacity), the neighbour links of
the forward step's zero-tile skip and the tile heights it runs on, and the
checkpoint resume decision (a checkpoint
with more shots done than the job has is refused)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gubingluo_amd")


def test_host_logic(tmp_path):
    lib = os.path.join(PKG, "librtm_hip.so")
    if not os.path.exists(lib):
        pytest.fail("librtm_hip.so is not built")
    exe = str(tmp_path / "host_logic")
    hipcc = "/opt/rocm/bin/hipcc"
    cmd = [hipcc, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc"),
           os.path.join(ROOT, "tests", "host_logic.cpp"), "-o", exe, "-L", PKG, "-lrtm_hip", f"-Wl,-rpath,{PKG}"]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stdout + b.stderr
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all host logic checks passed" in r.stdout
