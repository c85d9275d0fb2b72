"""Benchmark: MC trial moves/s of the device protonation / ion-swap chain.

Workload (default): the 30k-particle diblock brush -- 400 chains x 50 beads on
a 20x20 lattice in a 40x40x150 sigma box (x3 z-extension, slab correction),
beads 1-25 permanently +1, beads 26-50 titratable (pKa 7.5, pH 7.0), 2,000
salt pairs, one counter-particle per charged/titratable bead (44,000 atoms
realised); SPME 64x64x256, order 5, alpha 0.9, rc 3.5, WCA LJ; FP64 plain
path (exact per-move Delta-E, full recompute every 1,000 trials).

A step is one block of --trials-per-step GC trials.  W warm-up steps, then K
timed steps.  Inputs are resident on the GPU before timing starts.  Prints one
JSON line.  --gpus above 1 is refused: the chain of one replica runs on one
GPU (replicas would shard across GPUs with no communication).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def chain_bytes_per_trial(rep) -> float:
    """Algorithmic bytes one trial must read (each byte once), FP64 path:
    two full neighbour rows (4 B index + 8 B coefficient + 8 B neighbour charge
    + 4 B neighbour type per entry), two p^3 potential stencils (8 B each) and
    the 3 stencil-pair mesh-kernel windows ((2p-1)^3 x 8 B each)."""
    import numpy as np
    s, p = rep.s, rep.p
    h, _ = rep.neigh.download()
    d_full = 2.0 * h[-1] / s.natoms
    P = p.pme_order
    return 2 * d_full * 24 + 2 * P ** 3 * 8 + 3 * (2 * P - 1) ** 3 * 8


def refresh_bytes(rep) -> float:
    """Algorithmic HBM bytes of one potential refresh (FP64): the k-space pass
    writes the Gk-point complex half spectrum and reads the Gk influence values;
    the inverse plane pass reads the half spectrum and reads + writes the G-point
    potential grid."""
    nx, ny, nz = rep.p.grid
    G, Gk = nx * ny * nz, (nx // 2 + 1) * ny * nz
    return (16 * Gk + 8 * Gk) + (16 * Gk + 16 * G)


def cpu_baseline(cfg, trials: int, threads: int):
    from oracle import pyoracle as O
    from ph_brush_amd import system as S
    s, p, mc = cfg()
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    O.set_threads(threads)
    o.nlist()
    o.total_energy()
    t0 = time.perf_counter()
    r = o.gc_run(mc, trials, 1, record=False)
    dt = time.perf_counter() - t0
    return {"value": r["n"] / dt, "unit": "trial moves/s", "cores": threads, "kind": "port",
            "sample": f"{r['n']} GC trials of the same 44,000-atom workload, seed 1, reference algorithm "
                      f"(full LJ + SPME recompute per trial, FP64, OpenMP) in {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--trials-per-step", type=int, default=20000)
    ap.add_argument("--pending-max", type=int, default=8)
    ap.add_argument("--spec-window", type=int, default=8)
    ap.add_argument("--cpu-trials", type=int, default=0, help="0 = auto (about 10-20 s of CPU work)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs one replica chain on one GPU "
              "(independent replicas would shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    from ph_brush_amd import system as S
    from ph_brush_amd import _lib as L

    if L.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    cfg = S.config_30k_fp64
    s, p, mc = cfg()
    t_setup = time.perf_counter()
    rep = S.Replica(s, p, mc, seed=1, pending_max=args.pending_max, recompute_interval=1000,
                    spec_window=args.spec_window)
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] setup {t_setup:.2f}s natoms={s.natoms}", file=sys.stderr, flush=True)

    K, W, T = args.steps, args.warmup, args.trials_per_step
    for w in range(W):
        rep.mc.run(T)
        print(f"[bench] warmup step {w + 1}/{W} done", file=sys.stderr, flush=True)
    st0 = rep.mc.stats()
    t0 = time.perf_counter()
    for k in range(K):
        rep.mc.run(T)
        print(f"[bench] step {k + 1}/{K} done", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    st1 = rep.mc.stats()
    trials = st1["attempted"] - st0["attempted"]
    value = trials / dt

    # live kernel timing with HIP events on the library stream (one extra step)
    rep.mc.set_profiling(True)
    stp0 = rep.mc.stats()
    rep.mc.run(T)
    stp1 = rep.mc.stats()
    rep.mc.set_profiling(False)
    launches = stp1["launches"] - stp0["launches"]
    refreshes = stp1["refreshes"] - stp0["refreshes"]
    ptrials = stp1["attempted"] - stp0["attempted"]
    chain_ms = stp1["chain_ms"]
    bpt = chain_bytes_per_trial(rep)
    avg_launch_s = chain_ms / 1e3 / max(launches, 1)
    achieved = bpt * (ptrials / max(launches, 1)) / avg_launch_s / 1e9
    rb = refresh_bytes(rep)
    refresh_gbs = rb / (stp1["refresh_ms"] / 1e3 / max(refreshes, 1)) / 1e9

    out = {
        "metric": "MC trial moves/s at 30k particles (plain FP64 GPU path)",
        "value": value,
        "unit": "trial moves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": dt / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "diblock brush 400x50 on 20x20 lattice, 40x40x150 box x3 slab, 2000 salt pairs, "
                        "SPME 64x64x256 order 5, fixed-mu GC protonation + ion-swap moves, pH 7.0 pKa 7.5",
            "natoms": s.natoms,
            "trials_per_step": T,
            "pending_max": args.pending_max,
            "spec_window": args.spec_window,
            "acceptance": (st1["accepted"] - st0["accepted"]) / max(trials, 1),
            "parallelism": "1 chain / 1 GPU",
        },
        "roofline": {
            "kernel": "k_mc_chain",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": avg_launch_s * 1e6,
            "bytes_per_trial": bpt,
            "trials_per_launch": ptrials / max(launches, 1),
            "note": "single-workgroup chain: latency-bound, see refresh for the bandwidth kernels",
        },
        "refresh": {
            "achieved_GBs": refresh_gbs,
            "frac": refresh_gbs / HBM_PEAK_GBS,
            "avg_us": stp1["refresh_ms"] * 1e3 / max(refreshes, 1),
            "bytes": rb,
            "share_of_time": stp1["refresh_ms"] / max(stp1["chain_ms"] + stp1["refresh_ms"], 1e-9),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        n = args.cpu_trials or 150
        print(f"[bench] cpu baseline: {n} trials on {threads} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, n, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
