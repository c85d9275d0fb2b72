"""Benchmark: MC trial moves/s at 30k particles, delayed-acceptance surrogate
form vs the plain exact form, on one GPU.

Workload: the 30k-particle diblock brush -- 400 chains x 50 beads on a 20x20
lattice in a 40x40x150 sigma box (x3 z-extension, slab correction), beads 1-25
permanently +1, beads 26-50 titratable (pKa 7.5, pH 7.0), 2,000 salt pairs, one
counter-particle per charged/titratable bead (44,000 atoms realised); SPME
64x64x256, order 5, alpha 0.9, rc 3.5, WCA LJ; FP64 exact Delta-E (full
recompute of energy and potential every 20,000 trials).

The forms run in the same process, each timed over K steps of
--trials-per-step GC trials after W warm-up steps:
  plain: every trial evaluates the exact Delta-E;
  da:    delayed acceptance, local-pair surrogate -- a physics part (frozen
         field plus exactly tracked pair sums) and a BF16-MFMA MLP screen each
         trial and only stage-1 acceptances get the exact Delta-E; the second
         stage keeps the exact chain's distribution, so every trial is an
         exact-equivalent trial;
  da_wphi: the same with the W . phi surrogate (one accepted move per launch).
Each surrogate is trained first (untimed, reported) on (descriptor, exact
Delta-E) pairs collected on the GPU from exact chains at pH 6-9.  `value` is
the local-pair DA form's trial moves/s; the plain form, both ratios (per trial
and per accepted move), effective sample sizes per second and alpha of every
form are reported beside it.  Inputs are resident on the GPU before timing.
--gpus above 1 is refused: one replica's chain runs on one GPU (replicas would
shard across GPUs with no communication).
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


def _traffic(kernel: str):
    """HBM bytes per launch of a kernel from the committed counter-pass record
    (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled by what the counter
    reads for a device copy of known size); None when there is no record."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f)[kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def chain_bytes_per_trial(rep, fp32: bool = False) -> float:
    """Algorithmic bytes one trial must read (each byte once): two full
    neighbour rows (4 B index, the real-space coefficient, 8 B r^2 for the LJ
    change, 8 B neighbour charge, 4 B neighbour type per entry), two p^3
    potential stencils and the 3 stencil-pair mesh-kernel windows ((2p-1)^3
    values each).  Coefficients, potential and mesh kernel are 8 B in the FP64
    form and 4 B in the FP32 form."""
    s, p = rep.s, rep.p
    h, _ = rep.neigh.download()
    d_full = 2.0 * h[-1] / s.natoms
    P = p.pme_order
    w = 4 if fp32 else 8
    return 2 * d_full * (4 + w + 8 + 8 + 4) + 2 * P ** 3 * w + 3 * (2 * P - 1) ** 3 * w


def integrated_autocorr(x, c: float = 5.0):
    """Integrated autocorrelation time tau = 1/2 + sum_{t>=1} rho(t) of a series
    (in samples), summed up to the first window M >= c tau(M) (Sokal's automatic
    window); the effective sample size of the series is len(x) / (2 tau).
    Returns (tau, window)."""
    import numpy as np
    x = np.asarray(x, dtype=np.float64)
    x = x - x.mean()
    n = len(x)
    if n < 4 or not np.any(x):
        return float("nan"), 0
    f = np.fft.rfft(x, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    acf /= acf[0]
    taus = 0.5 + np.cumsum(acf[1:])
    for m in range(1, n):
        if m >= c * taus[m - 1]:
            return float(taus[m - 1]), m
    return float(taus[-1]), n - 1


def ess_of_chain(nprot0, E0, fwd, acc, dE, rate, label):
    """Effective samples of N_prot and of E per trial and per second, from a
    chain's per-trial record (direction, accepted flag, exact Delta-E of the
    accepted trials) and its timed trial rate."""
    import numpy as np
    acc = np.asarray(acc) != 0
    step = np.where(acc, np.where(np.asarray(fwd) != 0, 1, -1), 0)
    nprot = nprot0 + np.cumsum(step)
    energy = E0 + np.cumsum(np.where(acc, np.nan_to_num(np.asarray(dE)), 0.0))
    out = {"trials": int(len(step))}
    for name, series in (("nprot", nprot), ("E", energy)):
        tau, m = integrated_autocorr(series)
        out[name] = {"tau_int_trials": tau, "window": m, "ess": len(series) / (2 * tau),
                     "ess_per_s": rate / (2 * tau)}
    print(f"[bench] {label} ESS: tau(N_prot) {out['nprot']['tau_int_trials']:.0f} trials, "
          f"tau(E) {out['E']['tau_int_trials']:.0f} trials", file=sys.stderr, flush=True)
    return out


def refresh_uses_direct(rep, pending_max: int) -> bool:
    """The library's automatic choice: the direct refresh where the grid's x
    extent is at most 64 and at most 16 changes are pending."""
    return rep.p.grid[0] <= 64 and pending_max <= 16


def refresh_bytes(rep, pending_max: int) -> float:
    """Algorithmic HBM bytes of one potential refresh (FP64).  Direct refresh:
    the G-point potential grid read and written once, the folded mesh kernel
    ((nx/2+1)(ny/2+1)(nz/2+1) values) read once.  K-space refresh: the k-space
    pass writes the Gk-point complex half spectrum and reads the Gk influence
    values; the inverse plane pass reads the half spectrum and reads + writes
    the potential grid."""
    nx, ny, nz = rep.p.grid
    G, Gk = nx * ny * nz, (nx // 2 + 1) * ny * nz
    if refresh_uses_direct(rep, pending_max):
        return 16 * G + 8 * (nx // 2 + 1) * (ny // 2 + 1) * (nz // 2 + 1)
    return (16 * Gk + 8 * Gk) + (16 * Gk + 16 * G)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, trials: int, threads: int, trials_one: int):
    """The oracle's restatement of the reference algorithm (a full LJ + SPME
    recompute per trial) on the host cores: `trials` trials on `threads`
    threads, then `trials_one` on one thread."""
    from oracle import pyoracle as O
    from ph_brush_amd import system as S
    s, p, mc = cfg()
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    o.nlist()

    def leg(n, nthreads):
        O.set_threads(nthreads)
        o.total_energy()
        t0 = time.perf_counter()
        r = o.gc_run(mc, n, 1, record=False)
        return r["n"], time.perf_counter() - t0

    n, dt = leg(trials, threads)
    out = {"value": n / dt, "unit": "trial moves/s", "cores": threads, "kind": "port",
           "sample": f"{n} GC trials of the same 44,000-atom workload, seed 1, reference algorithm "
                     f"(full LJ + SPME recompute per trial, FP64, OpenMP) in {dt:.1f} s",
           "cpu_model": _cpu_model(), "host_cpus": os.cpu_count()}
    if trials_one > 0:
        n1, dt1 = leg(trials_one, 1)
        out["one_thread"] = {"value": n1 / dt1, "sample": f"{n1} trials of the same chain on 1 thread in {dt1:.1f} s"}
    return out


def _timed(run, K, W, T, label):
    for w in range(W):
        run(T)
        print(f"[bench] {label} warmup step {w + 1}/{W} done", file=sys.stderr, flush=True)
    alphas = []
    t0 = time.perf_counter()
    for k in range(K):
        st = run(T)
        alphas.append(st["nprot"] / st["nsites"])
        print(f"[bench] {label} step {k + 1}/{K} done", file=sys.stderr, flush=True)
    return time.perf_counter() - t0, alphas


def fp32_form(S, L, SG, args):
    """The 2,640-atom brush in the FP32 form vs the FP64 form: trial moves/s of
    both chains (recompute every 1,000 trials), Delta-E error over 10^5 sampled
    moves against the FP64 path (which equals the oracle's full recompute within
    1e-10), total-energy error of the FP32 PME plan against the FP64 oracle."""
    import numpy as np
    s, p, mc = S.config_small_fp64()
    res = {}
    for name, prec in (("fp64", L.PHB_FP64), ("fp32", L.PHB_FP32)):
        rep = S.Replica(s, p, mc, seed=1, pending_max=2, spec_window=8, recompute_interval=1000, precision=prec)
        rep.mc.run(args.fp32_trials)            # warm-up from the start state
        t0 = time.perf_counter()
        rep.mc.run(args.fp32_trials)
        res[name] = (args.fp32_trials / (time.perf_counter() - t0), rep)
    _, t = res["fp64"][1].atoms.identity()
    rng = np.random.default_rng(3)
    i, j, f = SG.sample_moves(rng, t, s.z, mc, 100000)
    # both forms evaluated in the same state (a fresh FP32 chain built on the FP64 chain's current identities)
    q64, t64 = res["fp64"][1].atoms.identity()
    r32 = S.Replica(S.System(s.x, s.y, s.z, q64, t64, s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data), p, mc, seed=1,
                    precision=L.PHB_FP32)
    d64 = res["fp64"][1].mc.delta_e(i, j, f)
    d32 = r32.mc.delta_e(i, j, f)
    err = d32 - d64
    from oracle import pyoracle as O
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    o.nlist()
    e64, _ = o.total_energy()
    atoms = S.DeviceAtoms(s)
    box = S.box_of(s)
    neigh = S.build_half_neighbors_gpu(atoms, box, S.DeviceCSR.from_host(*S.bond_adjacency(s.natoms, s.bonds)), p.rc)
    params = S.DeviceParams(p)
    e32, _ = S.total_energy(atoms, params, neigh, S.PMEPlan(params, box, p.rc, precision=L.PHB_FP32), box)
    print(f"[bench] fp32 form: {res['fp32'][0]:.0f} vs fp64 {res['fp64'][0]:.0f} trial moves/s", file=sys.stderr,
          flush=True)
    return {"value": res["fp32"][0], "fp64_value": res["fp64"][0], "unit": "trial moves/s",
            "workload": "64x20 all-titratable brush, 2,640 atoms, SPME 32x32x128 order 5, pH 7.5",
            "delta_e_error_vs_fp64": {"moves": len(i), "rms_kT": float(np.sqrt(np.mean(err ** 2))),
                                      "max_abs_kT": float(np.abs(err).max()),
                                      "rms_delta_e_kT": float(np.sqrt(np.mean(d64 ** 2)))},
            "total_energy_rel_error_vs_oracle": abs(e32 - e64) / abs(e64)}


def learned_form(S, SG, args):
    """The learned surrogate on the 2,640-atom brush (64 chains x 20 titratable
    beads): trained on about 2x10^6 (descriptor, exact Delta-E) pairs from exact
    chains at pH 6-9 with the W . phi physics part (where the network carries the
    real-space energy of the environment).  At pH 6, 7.5 and 9, from a state
    equilibrated by the plain chain: alpha and trial / accepted moves/s of the
    plain chain, of delayed acceptance with the trained network, of delayed
    acceptance with a network that outputs zero (the method without it), and of
    the local-pair form with its own network (same number of pairs)."""
    import math
    import numpy as np
    s, p, _ = S.config_small_fp64()
    n_per_ph = max(4096, (args.learned_pairs // 7 // 4096) * 4096)
    spec = SG.CollectSpec(samples_per_ph=n_per_ph, per_state=4096, equilibrate=20000, between=2000)
    sur = S.Surrogate(len(spec.ph_values) * n_per_ph)
    col = SG.collect(s, p, sur, spec)
    tr = SG.train_holdout(sur, epochs=args.learned_epochs, batch=4096, lr=2e-3)
    spec_l = SG.CollectSpec(samples_per_ph=n_per_ph, per_state=4096, equilibrate=20000, between=2000,
                            local_pairs=True)
    sur_l = S.Surrogate(len(spec_l.ph_values) * n_per_ph)
    col_l = SG.collect(s, p, sur_l, spec_l)
    tr_l = SG.train_holdout(sur_l, epochs=args.learned_epochs, batch=4096, lr=2e-3)
    zero = S.Surrogate(0)
    theta = np.zeros_like(zero.get_params())
    theta[-1] = 1.0
    zero.set_params(theta)
    blocks, bt = 10, args.learned_trials // 10

    def sample(run, stats_of):
        a = []
        s0 = stats_of()
        t0 = time.perf_counter()
        for _ in range(blocks):
            st = run(bt)
            a.append(st["nprot"] / st["nsites"])
        dt = time.perf_counter() - t0
        s1 = stats_of()
        acc = (s1["accepted"] - s0["accepted"]) / max(s1["attempted"] - s0["attempted"], 1)
        st1 = (s1.get("stage1_accepted", 0) - s0.get("stage1_accepted", 0)) / max(s1["attempted"] - s0["attempted"], 1)
        m = sum(a) / len(a)
        sem = math.sqrt(sum((x - m) ** 2 for x in a) / (len(a) - 1) / len(a))
        return {"alpha": m, "sem": sem, "trials_per_s": blocks * bt / dt, "acceptance": acc,
                "accepted_per_s": blocks * bt * acc / dt, "stage1_acceptance": st1 if st1 > 0 else None}

    rows = []
    for ph in (6.0, 7.5, 9.0):
        mc = S.McParams.at_ph(ph)
        rep = S.Replica(s, p, mc, seed=1, pending_max=2, spec_window=8, recompute_interval=20000)
        rep.mc.run(args.learned_burn)
        q, ty = rep.atoms.identity()
        s_eq = S.System(s.x, s.y, s.z, q, ty, s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)
        row = {"pH": ph, "plain": sample(rep.mc.run, rep.mc.stats)}
        del rep
        for name, net, local in (("da", sur, False), ("da_zero_network", zero, False), ("da_local_pairs", sur_l, True)):
            da = S.Replica(s_eq, p, mc, seed=2, pending_max=8, recompute_interval=20000)
            da.mc.da_attach(r_desc=spec.r_desc, local_pairs=local)
            r = sample(lambda n: da.mc.da_run(net, n), lambda: da.mc.da_run(net, 0))
            r["stage2_acceptance"] = r["acceptance"] / max(r["stage1_acceptance"] or 1e-12, 1e-12)
            r["over_plain"] = r["trials_per_s"] / row["plain"]["trials_per_s"]
            r["over_plain_accepted"] = r["accepted_per_s"] / max(row["plain"]["accepted_per_s"], 1e-12)
            r["alpha_diff"] = r["alpha"] - row["plain"]["alpha"]
            row[name] = r
            del da
        rows.append(row)
        print(f"[bench] learned surrogate, 2,640 atoms, pH {ph}: alpha plain {row['plain']['alpha']:.4f}, "
              f"DA {row['da']['alpha']:.4f}; DA {row['da']['over_plain']:.2f}x trials, "
              f"{row['da']['over_plain_accepted']:.2f}x accepted (zero network "
              f"{row['da_zero_network']['over_plain_accepted']:.2f}x; local pairs "
              f"{row['da_local_pairs']['over_plain_accepted']:.2f}x)", file=sys.stderr, flush=True)
    return {"workload": "64x20 all-titratable brush, 2,640 atoms, SPME 32x32x128 order 5, W . phi surrogate",
            "training_pairs": col["samples"], "collect_s": col["seconds"], "train_s": tr["ms"] / 1e3,
            "rmse_holdout_kT": tr["rmse_holdout"], "mae_holdout_kT": tr["mae_holdout"],
            "target_std_kT": tr["target_std"], "trials_per_block": bt, "blocks": blocks,
            "local_pairs_network": {"training_pairs": col_l["samples"], "rmse_holdout_kT": tr_l["rmse_holdout"],
                                    "target_std_kT": tr_l["target_std"]},
            "note": "candidates are screened 64 per round; a launch of this form ends at its accepted move",
            "rows": rows}


def fp32_44k_form(S, L, s_eq, p, mc, kw, K, T, fp64_value, fp64_bpt):
    """The FP32 form on the benchmark's own workload (44,000 atoms, from the same
    equilibrated state): trial moves/s over K steps of T trials, and the live
    launch time of its chain kernel against its algorithmic bytes."""
    rep = S.Replica(s_eq, p, mc, seed=1, precision=L.PHB_FP32, **kw)
    rep.mc.run(T)
    t0 = time.perf_counter()
    for _ in range(K):
        rep.mc.run(T)
    dt = time.perf_counter() - t0
    rep.mc.set_profiling(True)
    a = rep.mc.stats()
    rep.mc.run(T)
    b = rep.mc.stats()
    rep.mc.set_profiling(False)
    launches = max(b["launches"] - a["launches"], 1)
    bpt = chain_bytes_per_trial(rep, fp32=True)
    launch_s = b["chain_ms"] / 1e3 / launches
    tpl = (b["attempted"] - a["attempted"]) / launches
    achieved = bpt * tpl / launch_s / 1e9
    value = K * T / dt
    print(f"[bench] fp32 form on the 44k brush: {value:.0f} trial moves/s", file=sys.stderr, flush=True)
    return {"value": value, "fp64_value": fp64_value, "over_fp64": value / fp64_value, "unit": "trial moves/s",
            "workload": "the benchmark's 44,000-atom diblock brush, same equilibrated state and seed",
            "roofline": {"kernel": "k_mc_chain (FP32 form)", "bound": "hbm", "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "avg_launch_us": launch_s * 1e6, "bytes_per_trial": bpt, "trials_per_launch": tpl,
                         "fp64_bytes_per_trial": fp64_bpt}}


def setup_form(S, reps: int = 5):
    """The set-up path on the 44,000-atom workload, each step timed on the host
    around synchronous calls: the half neighbour list (cells, counts, scan, fill,
    row sort) against its algorithmic bytes (FP64 positions 24 N, cell index and
    cell-sorted atom ids 8 N, heads 4 (N + 1), entries 4 nnz written), and the
    MC object's set-up (full rows, coefficients, stencils, mesh kernel, potential)."""
    s, p, mc = S.config_30k_fp64()
    atoms, box = S.DeviceAtoms(s), S.box_of(s)
    bonds = S.DeviceCSR.from_host(*S.bond_adjacency(s.natoms, s.bonds))
    S.build_half_neighbors_gpu(atoms, box, bonds, p.rc)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        neigh = S.build_half_neighbors_gpu(atoms, box, bonds, p.rc)
        ts.append(time.perf_counter() - t0)
    h, _ = neigh.download()
    nnz, N = int(h[-1]), s.natoms
    nb_bytes = 24 * N + 8 * N + 4 * (N + 1) + 4 * nnz
    params = S.DeviceParams(p)
    pme = S.PMEPlan(params, box, p.rc)
    tm = []
    for _ in range(2):
        t0 = time.perf_counter()
        m = S.MonteCarlo(atoms, params, neigh, pme, box, mc, 1)
        tm.append(time.perf_counter() - t0)
        del m
    best = min(ts)
    return {"natoms": N, "nnz_half": nnz, "nlist_ms": best * 1e3, "nlist_bytes": nb_bytes,
            "nlist_GBs": nb_bytes / best / 1e9, "nlist_frac": nb_bytes / best / 1e9 / HBM_PEAK_GBS,
            "mc_create_ms": min(tm) * 1e3,
            "note": "host wall time of synchronous calls (allocation included); kernel times in the rocprof summary"}


def program_form():
    """The command-line program (ph_brush_amd/bin/MC, the reference's `MC in.data
    params.in`) end to end on the 44,000-atom workload written as input files:
    file reading, set-up, phase 1 to 60 % protonation, phase 2 to patience 3000,
    three data files and report.txt written.  Wall time of the child process."""
    import re
    import subprocess
    import tempfile
    from ph_brush_amd import formats as F
    exe = Path(__file__).resolve().parent / "ph_brush_amd" / "bin" / "MC"
    if not exe.exists():
        raise RuntimeError(f"{exe} is missing: build with `python -m ph_brush_amd.build`")
    with tempfile.TemporaryDirectory() as tmp:
        data, prm = F.write_workload("diblock_44000", tmp)
        print("[bench] program: MC in.data params.in on the 44,000-atom workload", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        r = subprocess.run([str(exe), str(data), str(prm)], cwd=tmp, capture_output=True, text=True, timeout=300)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"MC failed ({r.returncode}): {r.stderr}")
        rep = (Path(tmp) / "report.txt").read_text()
    updates = re.findall(r"\[sgmc\] mu-update#\d+ .*? attempts=(\d+)/", rep)
    reached = re.search(r"\[sgmc\] reached target in epoch=\d+ step_in_block=(\d+)", rep)
    chunks_last_epoch = re.findall(r"\[sgmc\] chunk acc=\d+/\d+ attempts=(\d+) ", rep.split("mu-update#")[-1])
    p1 = sum(map(int, updates)) + (int(chunks_last_epoch[-1]) if chunks_last_epoch else 0) \
        + (int(reached.group(1)) if reached else 0)
    steps = re.findall(r"\[phase2\] step=(\d+) ", rep)
    return {"wall_s": wall, "atoms": 44000, "phase1_trials": p1,
            "phase2_trials_at_last_report_line": int(steps[-1]) if steps else 0,
            "note": "one process: read files, neighbour list, energies, phase 1 (mu control to 60 %), phase 2 "
                    "(4-point moves to patience 3000), write phase1/phase2 data and report.txt"}


def batched_form(S, args):
    """32 replicas of a pH x salt titration grid (uneven sizes and grids)
    advanced together; aggregate trial moves/s over timed sweeps after the
    autotuner has picked the form (graph replays or one persistent launch per
    sweep) and its shape; alpha per replica averaged over the timed sweeps; the
    dominant kernel's roofline; per-replica rates against serial runs of the
    same replicas (the same chains: a serial copy run for the same number of
    trials reaches the same state)."""
    import numpy as np
    t0 = time.perf_counter()
    grid = S.config_titration_grid()
    reps = [S.Replica(s, p, mc, seed=7 + k, recompute_interval=0) for k, (s, p, mc, _) in enumerate(grid)]
    b = S.Batch([r.mc for r in reps])
    t_setup = time.perf_counter() - t0
    print(f"[bench] batched: {len(reps)} replicas, {sum(r.s.natoms for r in reps)} atoms, setup {t_setup:.1f}s",
          file=sys.stderr, flush=True)
    tune = b.autotune(1)
    print(f"[bench] batched: autotune -> persistent={tune['persistent']} W={tune['spec_window']} "
          f"pending={tune['pending_max']} pairs={tune['graph_pairs']} launch_cap={tune['launch_cap']} "
          f"refresh tile={tune['refresh_ty']}x{tune['refresh_tz']}", file=sys.stderr, flush=True)
    b.run(args.batch_warmup)
    st0 = [r.mc.stats() for r in reps]
    nsw, ms, launches, replays = args.batch_sweeps, 0.0, 0, 0
    alpha = np.zeros((nsw, len(reps)))
    for k in range(nsw):
        st = b.run(1)
        ms += st["ms"]
        launches += st["launches"]
        replays += st["replays"]
        for q, r in enumerate(reps):
            ss = r.mc.stats()
            alpha[k, q] = ss["nprot"] / ss["nsites"]
    st1 = [r.mc.stats() for r in reps]
    trials = sum(b1["attempted"] - a1["attempted"] for a1, b1 in zip(st0, st1))
    rate = trials / (ms / 1e3)
    print(f"[bench] batched: {rate:.0f} trial moves/s aggregate", file=sys.stderr, flush=True)
    # roofline of the dominant kernel: every trial's algorithmic bytes and every refresh's
    pm = tune["pending_max"]
    byts = 0.0
    for r, a1, b1 in zip(reps, st0, st1):
        byts += (b1["attempted"] - a1["attempted"]) * chain_bytes_per_trial(r)
        byts += (b1["accepted"] - a1["accepted"]) / max(pm // 2, 1) * refresh_bytes(r, pm)
    achieved = byts / (ms / 1e3) / 1e9
    kern = "k_sweep_persistent" if tune["persistent"] else "k_chain_batch + refresh launches"
    per_rep = None
    if tune["persistent"]:
        run_ms, ref_ms = b.replica_times()
        nat = [g[0].natoms for g in grid]
        order = sorted(range(len(grid)), key=lambda k: nat[k])
        per_rep = {"run_ms_max": float(run_ms.max()), "run_ms_mean": float(run_ms.mean()),
                   "refresh_share": float(ref_ms.sum() / max(run_ms.sum(), 1e-9)), "replicas": []}
        for k in (order[0], order[len(order) // 2], order[-1]):
            s, p, mc, lab = grid[k]
            n_run = st1[k]["attempted"]
            ns = st1[k]["nsites"]
            r = S.Replica(s, p, mc, seed=7 + k, recompute_interval=0, pending_max=pm,
                          spec_window=tune["spec_window"])
            r.mc.run(n_run - ns)
            tq = time.perf_counter()
            r.mc.run(ns)
            dts = time.perf_counter() - tq
            same = r.mc.stats()["nprot"] == st1[k]["nprot"] and r.mc.stats()["accepted"] == st1[k]["accepted"]
            per_rep["replicas"].append({"label": lab, "natoms": nat[k], "sites": ns,
                                        "batched_per_s": ns / (run_ms[k] / 1e3), "serial_per_s": ns / dts,
                                        "ratio": (ns / (run_ms[k] / 1e3)) / (ns / dts),
                                        "serial_state_equal": bool(same)})
            del r
        print(f"[bench] batched: per-replica rates {per_rep['replicas']}", file=sys.stderr, flush=True)
    alphas = {lab.split(":")[0]: float(alpha[:, q].mean()) for q, (s, p, mc, lab) in enumerate(grid)}
    return {"value": rate, "unit": "trial moves/s (aggregate over replicas)", "replicas": len(reps),
            "atoms": sum(r.s.natoms for r in reps), "sweeps_timed": nsw, "ms": ms,
            "trials_timed": trials, "launches": launches, "graph_replays": replays,
            "autotuned": {"persistent": tune["persistent"], "spec_window": tune["spec_window"],
                          "pending_max": tune["pending_max"], "graph_pairs": tune["graph_pairs"],
                          "launch_cap": tune["launch_cap"], "trials_per_s": tune["trials_per_s"],
                          "refresh_tile": [tune["refresh_ty"], tune["refresh_tz"]]},
            "roofline": {"kernel": kern, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_sweep": byts / nsw,
                         "note": "algorithmic bytes: each trial's rows, stencils and kernel windows, and each "
                                 "refresh's grid pass, over the timed sweeps' wall time"},
            "per_replica": per_rep,
            "alpha_by_replica": alphas,
            "note": f"alpha: mean over the {nsw} timed sweeps after {args.batch_warmup} warm-up sweeps and the "
                    "autotuner's sweeps from p_prot = 0.5 starts; a serial copy of a replica run for the same "
                    "trials reaches the same state (serial_state_equal)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--equil-trials", type=int, default=1000000,
                    help="plain-chain trials run once (untimed) before either form; both start from that state")
    ap.add_argument("--trials-per-step", type=int, default=200000,
                    help="trials per step; the workload's 10^6 trial moves are 5 steps")
    # the plain chain's fastest settings on this workload (measured sweep: a refresh after every
    # accepted move keeps pending windows out of every trial's Delta-E)
    ap.add_argument("--pending-max", type=int, default=2)
    ap.add_argument("--spec-window", type=int, default=16)
    ap.add_argument("--recompute-interval", type=int, default=20000,
                    help="trials between full recomputes of the energy and potential (drift reset)")
    ap.add_argument("--da-samples-per-ph", type=int, default=131072)
    ap.add_argument("--da-epochs", type=int, default=60)
    ap.add_argument("--da-local-samples-per-ph", type=int, default=16384,
                    help="training samples per pH of the local-pair form's network (its residual target is "
                         "small, so fewer samples than the W.phi form's)")
    ap.add_argument("--cpu-trials", type=int, default=0, help="0 = auto (about 10-20 s of CPU work)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-program", action="store_true", help="skip timing the MC command-line program")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--train-only", action="store_true",
                    help="collect the surrogate's data on the 44k workload, train it, report the training rate and exit")
    ap.add_argument("--setup-only", action="store_true",
                    help="time the set-up path (neighbour list, MC object) on the 44k workload and exit")
    ap.add_argument("--da-real-space", action="store_true",
                    help="also run the W . phi form whose surrogate includes the real-space potential")
    ap.add_argument("--no-da-wphi", action="store_true", help="skip the W . phi form of the surrogate")
    ap.add_argument("--da-only", action="store_true",
                    help="the plain and local-pair forms only (no W . phi, FP32, batched, program, set-up or CPU "
                         "legs): a short command for kernel profiles")
    ap.add_argument("--da-pending-max", type=int, default=8,
                    help="pending budget of the local-pair delayed-acceptance chain (accepted moves per launch x 2)")
    ap.add_argument("--fp32-trials", type=int, default=100000)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-surrogate form on the 2,640-atom brush")
    ap.add_argument("--learned-pairs", type=int, default=2000000,
                    help="training pairs of the learned-surrogate form (pH 6-9, 2,640-atom brush)")
    ap.add_argument("--learned-epochs", type=int, default=40)
    ap.add_argument("--learned-burn", type=int, default=200000)
    ap.add_argument("--learned-trials", type=int, default=500000, help="trials per chain and pH (10 blocks)")
    ap.add_argument("--ess-trials", type=int, default=6000000,
                    help="recorded trials per form for the autocorrelation times of N_prot and E (0: skip)")
    ap.add_argument("--batch-warmup", type=int, default=2, help="untimed sweeps after autotuning")
    ap.add_argument("--batch-sweeps", type=int, default=5, help="timed sweeps of the batched replicas")
    args = ap.parse_args()
    if args.da_only:
        args.no_da_wphi = args.no_fp32 = args.no_batched = args.no_program = args.no_cpu_baseline = True
        args.no_learned = True
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs one replica chain on one GPU "
              "(independent replicas would shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    from ph_brush_amd import system as S
    from ph_brush_amd import surrogate as SG
    from ph_brush_amd import _lib as L

    if L.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.setup_only:
        print(json.dumps({"setup": setup_form(S)}))
        return
    if args.train_only:
        s, p, mc = S.config_30k_fp64()
        spec = SG.CollectSpec(samples_per_ph=args.da_samples_per_ph, per_state=4096, equilibrate=20000, between=2000)
        sur = S.Surrogate(len(spec.ph_values) * spec.samples_per_ph)
        col = SG.collect(s, p, sur, spec)
        tr = SG.train_holdout(sur, epochs=args.da_epochs, batch=4096, lr=2e-3)
        flop = 3 * 2 * (48 * 128 + 128 * 128 + 128)
        tflops = tr["steps"] * 4096 * flop / (tr["ms"] / 1e3) / 1e12
        print(json.dumps({"train": {"samples": col["samples"], "collect_s": col["seconds"], "steps": tr["steps"],
                                    "ms": tr["ms"], "us_per_step": tr["ms"] * 1e3 / max(tr["steps"], 1),
                                    "tflops": tflops, "mfma_frac": tflops / 2500.0,
                                    "rmse_holdout_kT": tr["rmse_holdout"], "target_std_kT": tr["target_std"]}}))
        return
    cfg = S.config_30k_fp64
    s, p, mc = cfg()
    K, W, T = args.steps, args.warmup, args.trials_per_step
    kw = dict(pending_max=args.pending_max, recompute_interval=args.recompute_interval, spec_window=args.spec_window)

    # ---- equilibration (untimed), then the plain form from there
    t_setup = time.perf_counter()
    rep = S.Replica(s, p, mc, seed=1, **kw)
    print(f"[bench] setup {time.perf_counter() - t_setup:.2f}s natoms={s.natoms}", file=sys.stderr, flush=True)
    t_eq = time.perf_counter()
    if args.equil_trials > 0:
        rep.mc.run(args.equil_trials)
    q_eq, ty_eq = rep.atoms.identity()
    s_eq = S.System(s.x, s.y, s.z, q_eq, ty_eq, s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)
    print(f"[bench] equilibrated {args.equil_trials} plain trials in {time.perf_counter() - t_eq:.1f}s "
          f"(alpha {rep.mc.stats()['nprot'] / rep.mc.stats()['nsites']:.4f})", file=sys.stderr, flush=True)
    st0 = rep.mc.stats()
    dt_plain, alpha_plain = _timed(rep.mc.run, K, W, T, "plain")
    st1 = rep.mc.stats()
    plain_trials = K * T
    plain_value = plain_trials / dt_plain
    plain_acc = (st1["accepted"] - st0["accepted"]) / max(st1["attempted"] - st0["attempted"], 1)

    # live kernel timing with HIP events on the library stream (one extra step)
    rep.mc.set_profiling(True)
    stp0 = rep.mc.stats()
    rep.mc.run(T)
    stp1 = rep.mc.stats()
    rep.mc.set_profiling(False)
    launches = stp1["launches"] - stp0["launches"]
    refreshes = stp1["refreshes"] - stp0["refreshes"]
    ptrials = stp1["attempted"] - stp0["attempted"]
    bpt = chain_bytes_per_trial(rep)
    avg_launch_s = stp1["chain_ms"] / 1e3 / max(launches, 1)
    achieved = bpt * (ptrials / max(launches, 1)) / avg_launch_s / 1e9
    rb = refresh_bytes(rep, args.pending_max)
    refresh_method = "direct" if refresh_uses_direct(rep, args.pending_max) else "kspace"
    refresh_gbs = rb / (stp1["refresh_ms"] / 1e3 / max(refreshes, 1)) / 1e9
    del rep
    ess = {}
    if args.ess_trials > 0:
        rep = S.Replica(s_eq, p, mc, seed=3, record_cap=args.ess_trials, **kw)
        st_e = rep.mc.stats()
        rep.mc.run(args.ess_trials)
        r = rep.mc.record(args.ess_trials)
        ess["plain"] = ess_of_chain(st_e["nprot"], st_e["E"], r["fwd"], r["acc"], r["dE"], plain_value, "plain")
        del rep, r

    # ---- surrogates: data from exact chains at pH 6-9, GPU training (untimed); one network per
    # physics part (it sets the residual the network learns)
    flop_fwd = 2 * (48 * 128 + 128 * 128 + 128)
    flop_train = 3 * flop_fwd

    def train_surrogate(local_pairs, real_space=False):
        t0 = time.perf_counter()
        n_ph = args.da_local_samples_per_ph if local_pairs else args.da_samples_per_ph
        spec = SG.CollectSpec(samples_per_ph=n_ph, per_state=4096, equilibrate=20000,
                              between=2000, local_pairs=local_pairs, real_space=real_space)
        sur = S.Surrogate(len(spec.ph_values) * spec.samples_per_ph)
        col = SG.collect(s_eq, p, sur, spec, log=lambda m: print(m, file=sys.stderr, flush=True))
        tr = SG.train_holdout(sur, epochs=args.da_epochs, batch=4096, lr=2e-3)
        print(f"[bench] surrogate (local pairs {local_pairs}, real space {real_space}): {col['samples']} samples, "
              f"holdout rmse {tr['rmse_holdout']:.3f} (target std {tr['target_std']:.3f}), "
              f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        return spec, sur, col, tr

    def da_form(spec, sur, label, **attach):
        """Times the delayed-acceptance chain (same workload, start state and seed as the plain
        form), then one profiled step for the live launch time, then an ESS run."""
        kw_da = dict(kw)
        kw_da["pending_max"] = args.da_pending_max if attach.get("local_pairs") else args.pending_max
        rep = S.Replica(s_eq, p, mc, seed=1, **kw_da)
        t_att = time.perf_counter()
        rep.mc.da_attach(r_desc=spec.r_desc, **attach)
        t_att = time.perf_counter() - t_att
        d0 = rep.mc.da_run(sur, 0)
        dt, alpha = _timed(lambda n: rep.mc.da_run(sur, n), K, W, T, label)
        d1 = rep.mc.da_run(sur, 0)
        n_att = max(d1["attempted"] - d0["attempted"], 1)
        v = K * T / dt
        st1 = (d1["stage1_accepted"] - d0["stage1_accepted"]) / n_att
        acc = (d1["accepted"] - d0["accepted"]) / n_att
        # launch times from HIP events alone (the kernel as timed), then the phase clocks
        rep.mc.set_profiling(True, phase_clock=False)
        q0 = rep.mc.da_run(sur, 0)
        q1 = rep.mc.da_run(sur, T)
        launch_s = (q1["chain_ms"] - q0["chain_ms"]) / 1e3 / max(q1["launches"] - q0["launches"], 1)
        rep.mc.set_profiling(True)
        p0 = rep.mc.da_run(sur, 0)
        p1 = rep.mc.da_run(sur, T)
        rep.mc.set_profiling(False)
        nl = max(p1["launches"] - p0["launches"], 1)
        tpl = (p1["attempted"] - p0["attempted"]) / nl
        acc_p = max(p1["accepted"] - p0["accepted"], 1)
        phases = ["prologue", "proposals", "descriptors_physics", "mlp", "stage1", "exact_dE", "stage2_updates",
                  "epilogue"]
        per_acc = {k: (p1["phase_us"][i] - p0["phase_us"][i]) / acc_p for i, k in enumerate(phases)}
        del rep
        form = {"value": v, "ms_per_step": dt / K * 1e3, "acceptance": acc, "stage1_acceptance": st1,
                "stage2_acceptance": acc / max(st1, 1e-12), "alpha": sum(alpha) / len(alpha),
                "over_plain": v / plain_value, "accepted_per_s": v * acc,
                "over_plain_accepted": v * acc / max(plain_value * plain_acc, 1e-12),
                "trials_per_round": n_att / max(d1["rounds"] - d0["rounds"], 1), "trials_per_launch": tpl,
                "accepted_per_launch": acc_p / nl, "launch_us": launch_s * 1e6,
                "refresh_us_per_accepted": (p1["refresh_ms"] - p0["refresh_ms"]) * 1e3 / acc_p,
                "chain_us_per_accepted_by_phase": per_acc, "attach_s": t_att,
                "pending_max": kw_da["pending_max"]}
        if args.ess_trials > 0:
            rep = S.Replica(s_eq, p, mc, seed=3, record_cap=args.ess_trials, **kw_da)
            rep.mc.da_attach(r_desc=spec.r_desc, **attach)
            st_e = rep.mc.stats()
            rep.mc.da_run(sur, args.ess_trials)
            r = rep.mc.da_record(args.ess_trials)
            e = ess_of_chain(st_e["nprot"], st_e["E"], r["fwd"], r["stage"] == 1, r["dE"], v, label)
            form["ess"] = e
            form["ess_over_plain"] = {k: e[k]["ess_per_s"] / ess["plain"][k]["ess_per_s"] for k in ("nprot", "E")}
            del rep, r
        return form

    spec, sur, col, tr = train_surrogate(local_pairs=True)
    train_tflops = tr["steps"] * 4096 * flop_train / (tr["ms"] / 1e3) / 1e12
    da = da_form(spec, sur, "da", local_pairs=True)
    value, dt_da = da["value"], da["ms_per_step"] * K / 1e3
    stage1 = da["stage1_acceptance"]
    # algorithmic bytes per delayed-acceptance trial: two 96-byte histogram rows, two 64-byte records,
    # two 8-byte T values, and the stage-1 share of an exact trial
    da_bpt = 2 * 48 * 2 + 2 * 64 + 2 * 8 + stage1 * bpt
    da_launch_s = da["launch_us"] / 1e6
    da_trials_per_launch = da["trials_per_launch"]
    da_achieved = da_bpt * da_trials_per_launch / da_launch_s / 1e9
    infer_tflops = flop_fwd * da_trials_per_launch / da_launch_s / 1e12
    # surrogate cost against the workload's 10^6 moves: time to train, trials after which DA has paid it back
    train_total_s = col["seconds"] + tr["ms"] / 1e3
    gain_per_trial_s = 1.0 / plain_value - 1.0 / value
    break_even = train_total_s / gain_per_trial_s if gain_per_trial_s > 0 else float("inf")
    ess["da"] = da.get("ess")
    if ess.get("da"):
        for k in ("nprot", "E"):
            ess[f"da_over_plain_{k}"] = da["ess_over_plain"][k]

    # ---- the W . phi form of the surrogate (one accepted move per launch), its own network
    da_wphi = None
    if not args.no_da_wphi:
        spec_w, sur_w, col_w, tr_w = train_surrogate(local_pairs=False)
        da_wphi = da_form(spec_w, sur_w, "da W.phi")
        da_wphi.update({"rmse_holdout_kT": tr_w["rmse_holdout"], "target_std_kT": tr_w["target_std"],
                        "collect_s": col_w["seconds"], "train_s": tr_w["ms"] / 1e3})
    da_rs = None
    if args.da_real_space:
        spec_r, sur_r, _, tr_r = train_surrogate(local_pairs=False, real_space=True)
        da_rs = da_form(spec_r, sur_r, "da W.phi real-space", real_space=True)
        da_rs.update({"rmse_holdout_kT": tr_r["rmse_holdout"], "target_std_kT": tr_r["target_std"]})

    out = {
        "metric": "MC trial moves/s at 30k particles; DA-surrogate vs plain at equal alpha(pH)",
        "value": value,
        "unit": "exact-equivalent trial moves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": dt_da / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "diblock brush 400x50 on 20x20 lattice, 40x40x150 box x3 slab, 2000 salt pairs, "
                        "SPME 64x64x256 order 5, fixed-mu GC protonation + ion-swap moves, pH 7.0 pKa 7.5",
            "form": "delayed acceptance, local-pair surrogate (BF16-MFMA MLP stage 1, exact FP64 stage 2)",
            "natoms": s.natoms,
            "trials_per_step": T,
            "equil_trials": args.equil_trials,
            "pending_max": args.pending_max,
            "da_pending_max": args.da_pending_max,
            "spec_window": args.spec_window,
            "parallelism": "1 chain / 1 GPU",
        },
        "forms": {
            "plain": {"value": plain_value, "ms_per_step": dt_plain / K * 1e3, "acceptance": plain_acc,
                      "alpha": sum(alpha_plain) / len(alpha_plain)},
            "da": da,
            "da_over_plain": value / plain_value,
            # accepted moves per second: a delayed-acceptance chain accepts a smaller share of its trials
            # than the plain Metropolis chain (stage 1 times stage 2), so per accepted move the gain differs
            "plain_accepted_per_s": plain_value * plain_acc,
            "da_accepted_per_s": da["accepted_per_s"],
            "da_over_plain_accepted": da["over_plain_accepted"],
            "da_wphi": da_wphi,
            "da_real_space": da_rs,
            "alpha_note": "alpha of the timed blocks (same start state and seed); equality at fixed pH is "
                          "tested over long runs in tests/test_gpu_da.py",
        },
        "ess": ess or None,
        "surrogate_cost": {"collect_s": col["seconds"], "train_s": tr["ms"] / 1e3, "total_s": train_total_s,
                           "break_even_trials": break_even,
                           "workload_1e6_moves_s": {"plain": 1e6 / plain_value,
                                                    "da_with_training": train_total_s + 1e6 / value,
                                                    "da_without_training": 1e6 / value}},
        "surrogate": {"samples": col["samples"], "collect_s": col["seconds"], "train_steps": tr["steps"],
                      "train_ms": tr["ms"], "rmse_holdout_kT": tr["rmse_holdout"],
                      "mae_holdout_kT": tr["mae_holdout"], "target_std_kT": tr["target_std"],
                      "train_tflops": train_tflops, "train_mfma_frac": train_tflops / 2500.0},
        "roofline": {
            "kernel": "k_da_chain",
            "bound": "hbm",
            "achieved": da_achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": da_achieved / HBM_PEAK_GBS,
            "traffic": _traffic("k_da_chain_local"),
            "traffic_unit": "HBM bytes per launch (FETCH_SIZE and WRITE_SIZE counter passes, copy-calibrated)",
            "avg_launch_us": da_launch_s * 1e6,
            "bytes_per_trial": da_bpt,
            "trials_per_launch": da_trials_per_launch,
            "mlp_inference_tflops": infer_tflops,
            "mlp_inference_mfma_frac": infer_tflops / 2500.0,
            "note": "single-workgroup chain: latency-bound; the MFMA fraction of the in-chain MLP is reported "
                    "against the dense BF16 peak",
        },
        "plain_roofline": {
            "kernel": "k_mc_chain",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "avg_launch_us": avg_launch_s * 1e6,
            "bytes_per_trial": bpt,
            "trials_per_launch": ptrials / max(launches, 1),
        },
        "refresh": {
            "method": refresh_method,
            "achieved_GBs": refresh_gbs,
            "frac": refresh_gbs / HBM_PEAK_GBS,
            "avg_us": stp1["refresh_ms"] * 1e3 / max(refreshes, 1),
            "bytes": rb,
            "share_of_time": stp1["refresh_ms"] / max(stp1["chain_ms"] + stp1["refresh_ms"], 1e-9),
        },
        "cpu_baseline": None,
    }
    if not args.no_fp32:
        out["forms"]["fp32"] = fp32_form(S, L, SG, args)
        out["forms"]["fp32_44k"] = fp32_44k_form(S, L, s_eq, p, mc, kw, K, T, plain_value, bpt)
    if not args.no_batched:
        out["forms"]["batched"] = batched_form(S, args)
    if not args.no_learned:
        out["forms"]["learned_2640"] = learned_form(S, SG, args)
    if not args.no_program:
        out["forms"]["program"] = program_form()
    if not args.da_only:
        out["setup"] = setup_form(S)
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        n = args.cpu_trials or 150
        print(f"[bench] cpu baseline: {n} trials on {threads} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg, n, threads, max(1, n // 12))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
