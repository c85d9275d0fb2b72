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
    # the trace, collected window by window (97 records at a time), covers every trial
    assert kv["trace_n"][0] == s2["attempted"] and kv["trace_acc"][0] == s2["accepted"]
    assert abs(kv["trace_E"][0] - s2["E"]) <= 1e-8 * max(1.0, abs(s2["E"]))


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


def _build_dropin(tmp_path):
    exe = tmp_path / "dropin_demo"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "cpp" / "dropin_demo.cpp"), f"-L{ROOT / 'ph_brush_amd'}", "-lphb",
           f"-Wl,-rpath,{ROOT / 'ph_brush_amd'}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_caller_compiles_against_reference_signatures(tmp_path):
    """A caller typed with the reference's move signatures (MC/include/mc.hpp:50-81,
    std::mt19937& and Params& included) compiles and, without a GPU, fails the way
    main does."""
    exe = _build_dropin(tmp_path)
    if L.device_count() > 0:
        pytest.skip("a GPU is visible")
    s, _, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2, n_salt=10)
    _write_system(s, tmp_path / "sys.txt")
    r = subprocess.run([str(exe), str(tmp_path / "sys.txt"), "75", "90"], capture_output=True, text=True)
    assert r.returncode == 2 and "Error:" in r.stderr and "no CPU fallback" in r.stderr


def _kv(text):
    kv = {}
    for line in text.splitlines():
        k, *v = line.split()
        kv[k] = [float(x) for x in v]
    return kv


@pytest.mark.gpu
@pytest.mark.parametrize("target,threshold", [(75.0, 90), (95.0, 90)])
def test_dropin_caller_matches_oracle(gpu, tmp_path, target, threshold):
    """main()'s sequence through the reference signatures against the oracle's
    phase 1 and phase 2 (MC/src/main.cu:180-206): counts, epochs, bit-identical
    mu written back into Params, identities written back into System, the
    caller's std::mt19937 after each phase equal to the oracle's engine, and
    phase 2 from the Phase1Result lists."""
    from oracle import pyoracle as O
    exe = _build_dropin(tmp_path)
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2, n_salt=10)
    _write_system(s, tmp_path / "sys.txt")
    out = subprocess.run([str(exe), str(tmp_path / "sys.txt"), str(target), str(threshold)], capture_output=True,
                         text=True, check=True)
    kv = _kv(out.stdout)
    mc = S.McParams(beta=1.0, mu_eff=0.0)
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    o.nlist()
    ro = o.phase1_run(mc, 5, target, direct_pct_threshold=threshold, gc_block=200, epoch_attempt_max=1000)
    assert ro["rc"] == 0 and kv["use3dc"] == [1.0]
    assert kv["direct"][0] == ro["direct"]
    assert (kv["p1_attempted"][0], kv["p1_accepted"][0]) == (ro["attempted"], ro["accepted"])
    assert abs(kv["p1_E"][0] - ro["E"]) <= 1e-9 * max(1.0, abs(ro["E"]))
    if not ro["direct"]:
        assert kv["blocks"][0] == ro["blocks_done"] and kv["reached"][0] == ro["reached_target"] == 1
        assert (kv["fwd"][0], kv["bwd"][0]) == (ro["accepted_forward"], ro["accepted_backward"])
        assert kv["mu_eff"][0] == ro["mu_eff"]                # Params& p.mu_eff, bit for bit
    assert kv["handover"] == [float(x) for x in ro["handover_sizes"]]
    assert np.array_equal(np.array(kv["types1"], np.int64), ro["type0"] + 1)
    assert np.array_equal(np.array(kv["engine1"], np.uint64), ro["rng_state"].astype(np.uint64))
    s1 = S.System(s.x, s.y, s.z, ro["q"], ro["type0"], s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)
    o2 = O.Oracle(s1, p, S.bond_adjacency(s.natoms, s.bonds))
    o2.nlist()
    r2 = o2.phase2_run(mc, ro["rng_state"], 600, 100000, pools=ro["handover"], pool_sizes=ro["handover_sizes"])
    assert (kv["p2_attempted"][0], kv["p2_accepted"][0]) == (r2["n"], r2["accepted"]) and r2["accepted"] > 0
    assert abs(kv["p2_best"][0] - r2["best_E"]) <= 1e-9 * max(1.0, abs(r2["best_E"]))
    assert np.array_equal(np.array(kv["types2"], np.int64), r2["type0"] + 1)
    assert np.array_equal(np.array(kv["engine2"], np.uint64), r2["rng_state"].astype(np.uint64))
