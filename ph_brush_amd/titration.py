"""alpha(pH) of a brush with the plain chain and the delayed-acceptance chain.

One surrogate is trained on exact chains at pH 6-9.  At each pH a plain chain
equilibrates from the workload's start; the plain chain and the delayed-
acceptance chain then both continue from that equilibrated state (same
identities), and alpha = N_prot / N_sites is sampled after every block of
trials.  Delayed acceptance keeps the exact chain's stationary distribution,
so the two curves must agree within their sampling error.  Each row also
reports the wall-time ratio at equal trials and the ratio of accepted moves per
second (a delayed-acceptance chain accepts a slightly smaller share of its
trials: stage 1 times stage 2).  The delayed-acceptance chain runs the
local-pair form by default (local_pairs=False: the W . phi form).

python -m ph_brush_amd.titration [--blocks 20 --block-trials 100000 ...]
prints one line per pH and a JSON summary.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

from . import surrogate as SG
from . import system as S


def _sample(run, blocks, block_trials):
    a = []
    for _ in range(blocks):
        st = run(block_trials)
        a.append(st["nprot"] / st["nsites"])
    mean = sum(a) / len(a)
    var = sum((x - mean) ** 2 for x in a) / max(len(a) - 1, 1)
    return mean, math.sqrt(var / len(a)), a


def titration(cfg=S.config_30k_fp64, ph_values=(6.0, 6.5, 7.0, 7.5, 8.0, 8.5, 9.0), burn=300000, blocks=20,
              block_trials=100000, samples_per_ph=0, epochs=60, seed=1, log=print, local_pairs=True):
    """samples_per_ph 0: 16,384 for the local-pair form (its network's residual
    target is small), 131,072 for the W . phi form."""
    s, p, _ = cfg()
    if samples_per_ph <= 0:
        samples_per_ph = 16384 if local_pairs else 131072
    t0 = time.perf_counter()
    spec = SG.CollectSpec(samples_per_ph=samples_per_ph, per_state=4096, equilibrate=20000, between=2000,
                          local_pairs=local_pairs)
    sur = S.Surrogate(len(spec.ph_values) * spec.samples_per_ph)
    SG.collect(s, p, sur, spec)
    tr = SG.train_holdout(sur, epochs=epochs, batch=4096, lr=2e-3)
    log(f"surrogate: holdout rmse {tr['rmse_holdout']:.3f} (target std {tr['target_std']:.3f}), "
        f"{time.perf_counter() - t0:.1f} s")
    rows = []
    for ph in ph_values:
        mc = S.McParams.at_ph(ph)
        rep = S.Replica(s, p, mc, seed=seed, pending_max=2, spec_window=16, recompute_interval=20000)
        rep.mc.run(burn)
        q, ty = rep.atoms.identity()
        s_eq = S.System(s.x, s.y, s.z, q, ty, s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)
        p0 = rep.mc.stats()
        tp = time.perf_counter()
        a_plain, e_plain, _ = _sample(rep.mc.run, blocks, block_trials)
        tp = time.perf_counter() - tp
        acc_p = rep.mc.stats()["accepted"] - p0["accepted"]
        del rep
        da = S.Replica(s_eq, p, mc, seed=seed + 1, pending_max=8, recompute_interval=20000)
        da.mc.da_attach(r_desc=spec.r_desc, local_pairs=local_pairs)
        d0 = da.mc.da_run(sur, 0)
        td = time.perf_counter()
        a_da, e_da, _ = _sample(lambda n: da.mc.da_run(sur, n), blocks, block_trials)
        td = time.perf_counter() - td
        acc_d = da.mc.da_run(sur, 0)["accepted"] - d0["accepted"]
        del da
        n = blocks * block_trials
        row = {"pH": ph, "alpha_plain": a_plain, "sem_plain": e_plain, "alpha_da": a_da, "sem_da": e_da,
               "diff": a_da - a_plain, "plain_s": tp, "da_s": td, "speedup": tp / td,
               "acceptance_plain": acc_p / n, "acceptance_da": acc_d / n,
               "speedup_accepted": (acc_d / td) / max(acc_p / tp, 1e-12)}
        rows.append(row)
        log(f"pH {ph:4.1f}: alpha plain {a_plain:.4f} +- {e_plain:.4f}, DA {a_da:.4f} +- {e_da:.4f}, "
            f"diff {a_da - a_plain:+.4f}; {n} trials each, plain {tp:.1f} s, DA {td:.1f} s "
            f"({tp / td:.2f}x trials, {row['speedup_accepted']:.2f}x accepted moves)")
    return {"workload": f"{s.natoms} atoms", "form": "local pairs" if local_pairs else "W . phi", "burn": burn,
            "blocks": blocks, "block_trials": block_trials,
            "max_abs_diff": max(abs(r["diff"]) for r in rows),
            "max_diff_in_sem": max(abs(r["diff"]) / max(math.hypot(r["sem_plain"], r["sem_da"]), 1e-12) for r in rows),
            "min_speedup": min(r["speedup"] for r in rows),
            "min_speedup_accepted": min(r["speedup_accepted"] for r in rows), "rows": rows}


def main():
    ap = argparse.ArgumentParser(description="alpha(pH) with the plain and the delayed-acceptance chains")
    ap.add_argument("--burn", type=int, default=300000)
    ap.add_argument("--blocks", type=int, default=20)
    ap.add_argument("--block-trials", type=int, default=100000)
    ap.add_argument("--samples-per-ph", type=int, default=0,
                    help="0: 16,384 (local pairs) or 131,072 (W . phi)")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--wphi", action="store_true", help="the W . phi form of the surrogate instead of local pairs")
    a = ap.parse_args()
    res = titration(burn=a.burn, blocks=a.blocks, block_trials=a.block_trials, samples_per_ph=a.samples_per_ph,
                    epochs=a.epochs, log=lambda m: print(m, file=sys.stderr, flush=True), local_pairs=not a.wphi)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
