"""Writes tests/golden/oracle_tiny.npz: golden vectors of the oracle (the FP64
CPU restatement of MC/src/neighbor.cu:74-125, MC/src/pme.cu:291-457,
MC/src/mc.cu:43-60 and the trial loop MC/src/mc.cu:379-503) on two small
brushes with PME and one with coulombtype coul_cut, with every input stored
beside the outputs:

  - system: positions, charges, 0-based types, bonds, box (data-file height);
  - parameters: cutoffs, epsilon_r, Ewald alpha, spline order, grid, slab flag;
  - phases: phase 1 (mu control) and phase 2 from its hand-over on one brush
    (counts, mu, identities, engine state, pools; phase-2 trials);
  - outputs: the half neighbour list (CSR), the total energy and its
    components, and a 400-trial fixed-mu GC chain (seed 5): direction, site,
    ion, decision, Delta-E, Metropolis exponent and uniform per trial, final identities
    and energy.

The oracle's energies are OpenMP sums, so a rerun reproduces the floating-point
vectors to about 1e-13 (the tests allow 1e-12) and the integer ones exactly.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402
from ph_brush_amd import _lib as L  # noqa: E402
from ph_brush_amd import system as S  # noqa: E402

CASES = [("o4", dict(order=4, grid=(16, 16, 48), n_perm=2, n_salt=8), 7.9, "pme"),
         ("o5", dict(order=5, grid=(16, 20, 45), n_perm=0, n_salt=6), 7.5, "pme"),
         ("c4", dict(order=4, grid=(16, 16, 48), n_perm=2, n_salt=8), 7.9, "coul_cut")]
NTRIALS, SEED = 400, 5


def main():
    out = {}
    meta = {}
    for name, kw, ph, coulomb in CASES:
        s, p, _ = S.config_tiny(**kw)
        if coulomb == "coul_cut":
            p.coulombtype = L.PHB_COUL_CUT
        mc = S.McParams.at_ph(ph)
        o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
        head, lst = o.nlist()
        e, comps = o.total_energy()
        r = o.gc_run(mc, NTRIALS, SEED)
        assert r["n"] == NTRIALS
        pre = name + "_"
        out.update({pre + "x": s.x, pre + "y": s.y, pre + "z": s.z, pre + "q": s.q, pre + "type0": s.type0,
                    pre + "bonds": s.bonds, pre + "head": head, pre + "list": lst,
                    pre + "E": np.array([e]), pre + "comps": comps,
                    pre + "fwd": r["fwd"], pre + "i": r["i"], pre + "j": r["j"], pre + "acc": r["acc"],
                    pre + "dE": r["dE"], pre + "expo": r["expo"], pre + "u": r["u"], pre + "q_end": r["q"],
                    pre + "type0_end": r["type0"], pre + "E_end": np.array([r["E"]])})
        meta[name] = {"box": [s.Lx, s.Ly, s.Lz, s.Lz_data],
                      "params": {"rc": p.rc, "lj_rc": p.lj_rc, "lj_shift": p.lj_shift, "epsilon_r": p.epsilon_r,
                                 "ewald_alpha": p.ewald_alpha, "pme_order": p.pme_order, "grid": list(p.grid),
                                 "pme_3dc": p.pme_3dc, "ntypes": p.ntypes, "coulombtype": p.coulombtype},
                      "mc": {"nh_n_type": mc.nh_n_type, "nh_p_type": mc.nh_p_type, "w_type": mc.w_type,
                             "cl_type": mc.cl_type, "beta": mc.beta, "mu_eff": mc.mu_eff, "w_z_extra": mc.w_z_extra},
                      "ntrials": NTRIALS, "seed": SEED}
    # phase 1 (mu control to 75 %, blocks of 200, epochs of at most 1000 trials, seed 5) and
    # phase 2 (600 trials from the hand-over, patience 10^6) on the o5 brush
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2, n_salt=10)
    mc = S.McParams.at_ph(7.0)
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    o.nlist()
    ro = o.phase1_run(mc, SEED, 75.0, gc_block=200, epoch_attempt_max=1000)
    assert ro["rc"] == 0 and ro["reached_target"] == 1
    s1 = S.System(s.x, s.y, s.z, ro["q"], ro["type0"], s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)
    o2 = O.Oracle(s1, p, S.bond_adjacency(s.natoms, s.bonds))
    o2.nlist()
    r2 = o2.phase2_run(mc, ro["rng_state"], 600, 10 ** 6, pools=ro["handover"], pool_sizes=ro["handover_sizes"])
    counts = [ro[k] for k in ("attempted", "accepted", "accepted_forward", "accepted_backward", "blocks_done",
                              "Nh", "Nprot", "target_count")]
    out.update({"p1_counts": np.array(counts, np.int64), "p1_mu": np.array([ro["mu_eff"]]),
                "p1_E": np.array([ro["E"]]), "p1_type0": ro["type0"], "p1_rng": ro["rng_state"],
                "p1_handover": ro["handover"], "p1_handover_sizes": ro["handover_sizes"],
                "p2_atoms": r2["atoms"], "p2_acc": r2["acc"], "p2_dE": r2["dE"], "p2_type0": r2["type0"],
                "p2_E": np.array([r2["E"]]), "p2_mode": np.array([r2["mode"]])})
    meta["phases"] = {"case": "config_tiny(order=5, grid=(16, 16, 48), n_perm=2, n_salt=10), pH 7.0",
                      "target_pct": 75.0, "gc_block": 200, "epoch_attempt_max": 1000, "seed": SEED,
                      "phase2_trials": 600, "patience": 10 ** 6}
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    path = os.path.join(ROOT, "tests", "golden", "oracle_tiny.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
