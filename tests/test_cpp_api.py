"""The reference-style C++ API (include/ph_brush/ph_brush.hpp) compiles against
the C ABI, links libphb.so, reports errors the way the reference's main does
(exit 2 with "Error: ..."), and on the GPU gives the same numbers as the
Python path over the same C ABI."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ph_brush_amd import _lib as L
from ph_brush_amd import system as S

ROOT = Path(__file__).resolve().parent.parent


def _build(tmp_path):
    exe = tmp_path / "api_demo"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "cpp" / "api_demo.cpp"), f"-L{ROOT / 'ph_brush_amd'}", "-lphb",
           f"-Wl,-rpath,{ROOT / 'ph_brush_amd'}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _write_system(s, path):
    with open(path, "w") as f:
        f.write(f"{s.natoms} {len(s.bonds)} {float(s.Lx)!r} {float(s.Ly)!r} {float(s.Lz)!r}\n")
        for i in range(s.natoms):
            f.write(f"{float(s.x[i])!r} {float(s.y[i])!r} {float(s.z[i])!r} {float(s.q[i])!r} {int(s.type0[i]) + 1}\n")
        for i, j in s.bonds:
            f.write(f"{i} {j}\n")


def test_cpp_api_compiles_and_fails_loudly_without_gpu(tmp_path):
    exe = _build(tmp_path)
    if L.device_count() > 0:
        pytest.skip("a GPU is visible")
    s, _, _ = S.config_tiny(order=5, grid=(16, 16, 48))
    _write_system(s, tmp_path / "sys.txt")
    r = subprocess.run([str(exe), str(tmp_path / "sys.txt"), "10"], capture_output=True, text=True)
    assert r.returncode == 2 and "Error:" in r.stderr and "no CPU fallback" in r.stderr


@pytest.mark.gpu
def test_cpp_api_matches_python_path(gpu, tmp_path):
    exe = _build(tmp_path)
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48))
    _write_system(s, tmp_path / "sys.txt")
    out = subprocess.run([str(exe), str(tmp_path / "sys.txt"), "500"], capture_output=True, text=True, check=True)
    kv = {}
    for line in out.stdout.splitlines():
        k, *v = line.split()
        kv[k] = [float(x) for x in v]
    mc = S.McParams(beta=1.0, mu_eff=0.0)
    rep = S.Replica(s, p, mc, seed=1)
    e, comps = rep.pme.energy(rep.atoms, rep.neigh, rep.box)
    assert kv["grid"] == [16, 16, 48, 5]
    assert abs(kv["E_coul"][0] - e) <= 1e-12 * max(1.0, abs(e))
    assert abs(kv["E_recip"][0] - comps.recip_space) <= 1e-12 * max(1.0, abs(comps.recip_space))
    st = rep.mc.run(500)
    assert kv["attempted"][0] == st["attempted"] == 500
    assert kv["accepted"][0] == st["accepted"] and kv["nprot"][0] == st["nprot"]
    assert abs(kv["E_after"][0] - st["E"]) <= 1e-9 * max(1.0, abs(st["E"]))
    rep.mc.phase2_begin(patience=200)
    s2 = rep.mc.phase2_run(400)
    assert kv["p2_attempted"][0] == s2["attempted"] and kv["p2_accepted"][0] == s2["accepted"]
    assert abs(kv["p2_best"][0] - s2["best_E"]) <= 1e-9 * max(1.0, abs(s2["best_E"]))


@pytest.mark.gpu
def test_cpp_api_phase1_then_phase2_matches_python_path(gpu, tmp_path):
    """main()'s sequence through the C++ layer: phase_gc_mu_control to a target,
    the hand-over, phase 2; the same calls from Python give the same numbers."""
    exe = _build(tmp_path)
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48))
    _write_system(s, tmp_path / "sys.txt")
    out = subprocess.run([str(exe), str(tmp_path / "sys.txt"), "100000", "75"], capture_output=True, text=True,
                         check=True)
    kv = {}
    for line in out.stdout.splitlines():
        k, *v = line.split()
        kv[k] = [float(x) for x in v]
    rep = S.Replica(s, p, S.McParams(beta=1.0, mu_eff=0.0), seed=1)
    g = rep.mc.phase1_run(75.0, direct_pct_threshold=1000, gc_block=200, epoch_attempt_max=1000,
                          max_trials=100000)
    ho = rep.mc.phase1_handover()
    assert g["reached_target"] == 1 and kv["reached"][0] == 1
    assert kv["attempted"][0] == g["attempted"] and kv["accepted"][0] == g["accepted"]
    assert kv["blocks"][0] == g["blocks_done"] and kv["nprot"][0] == g["Nprot"] == g["target_count"]
    assert kv["mu"][0] == g["mu_eff"]
    assert kv["handover"] == [len(ho[k]) for k in ("nh_n_remaining", "nh_p_mc", "w_remaining", "cl_mc")]
    assert len(ho["nh_p_mc"]) == g["target_count"]
    rep.mc.phase2_begin(patience=200)
    s2 = rep.mc.phase2_run(400)
    assert kv["p2_attempted"][0] == s2["attempted"] and kv["p2_accepted"][0] == s2["accepted"]
    assert abs(kv["p2_best"][0] - s2["best_E"]) <= 1e-9 * max(1.0, abs(s2["best_E"]))
