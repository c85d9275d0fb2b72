"""Golden vectors (tests/golden/oracle_tiny.npz, written by
tests/golden/make_golden.py from the oracle): the half neighbour list
(MC/src/neighbor.cu:74-125), the total energy and its components
(MC/src/mc.cu:43-60, MC/src/pme.cu:344-457) and a 400-trial fixed-mu GC chain
(MC/src/mc.cu:379-503) on two small brushes, inputs stored beside outputs.

CPU: the oracle reproduces them (a regression pin of the restatement).
GPU: the device path matches them to the parity bar -- identical list,
energies within 1e-10 relative, identical accept/reject sequence (a split is
allowed only within 1e-9 of the Metropolis threshold), Delta-E within 1e-10.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import _lib as L
from ph_brush_amd import system as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_tiny.npz")


def _load():
    z = np.load(GOLDEN)   # allow_pickle stays False
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


def _case(z, meta, name):
    m = meta[name]
    Lx, Ly, Lz, Lzd = m["box"]
    g = lambda k: z[name + "_" + k]  # noqa: E731
    s = S.System(g("x"), g("y"), g("z"), g("q"), g("type0").astype(np.int32), g("bonds").astype(np.int32),
                 Lx, Ly, Lz, Lzd)
    pp = m["params"]
    p = S.Params(rc=pp["rc"], lj_rc=pp["lj_rc"], lj_shift=pp["lj_shift"], epsilon_r=pp["epsilon_r"],
                 ewald_alpha=pp["ewald_alpha"], pme_order=pp["pme_order"], grid=tuple(pp["grid"]),
                 pme_3dc=pp["pme_3dc"], ntypes=pp["ntypes"], coulombtype=pp.get("coulombtype", L.PHB_COUL_PME))
    mc = S.McParams(**m["mc"])
    return s, p, mc, g


CASES = ["o4", "o5", "c4"]   # c4: coulombtype coul_cut


def _rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden_vectors(name):
    z, meta = _load()
    s, p, mc, g = _case(z, meta, name)
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    head, lst = o.nlist()
    assert np.array_equal(head, g("head")) and np.array_equal(lst, g("list"))
    e, comps = o.total_energy()
    assert _rel(e, g("E")[0]) <= 1e-13
    assert np.all(np.abs(comps - g("comps")) <= 1e-13 * np.maximum(1.0, np.abs(g("comps"))))
    r = o.gc_run(mc, meta[name]["ntrials"], meta[name]["seed"])
    for k in ("fwd", "i", "j", "acc"):
        assert np.array_equal(r[k], g(k)), k
    assert np.all(np.abs(r["dE"] - g("dE")) <= 1e-12 * np.maximum(1.0, np.abs(g("dE"))))
    assert np.array_equal(r["type0"], g("type0_end")) and _rel(r["E"], g("E_end")[0]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_device_matches_golden_vectors(gpu, name):
    z, meta = _load()
    s, p, mc, g = _case(z, meta, name)
    n, seed = meta[name]["ntrials"], meta[name]["seed"]
    r = S.Replica(s, p, mc, seed=seed, pending_max=8, recompute_interval=100, record_cap=n)
    head, lst = r.neigh.download()
    assert np.array_equal(head, g("head")) and np.array_equal(lst, g("list"))
    e, comps = r.total_energy()
    assert _rel(e, g("E")[0]) <= 1e-10
    assert np.all(np.abs(comps - g("comps")) <= 1e-10 * np.maximum(1.0, np.abs(g("E")[0])))
    st = r.mc.run(n)
    assert st["attempted"] == n and st["status"] == 0
    rec = r.mc.record(n)
    k = n
    for t in range(n):
        if not (rec["fwd"][t] == g("fwd")[t] and rec["i"][t] == g("i")[t] and rec["j"][t] == g("j")[t]):
            raise AssertionError(f"proposal diverged at trial {t} without a prior decision split")
        if rec["acc"][t] != g("acc")[t]:
            u = g("u")[t]
            thr = 0.0 if math.isnan(u) else math.log(u)
            assert abs(g("expo")[t] - thr) < 1e-9 or abs(g("expo")[t]) < 1e-9, f"decision split at trial {t}"
            k = t
            break
    dE = g("dE")[:k]
    assert np.all(np.abs(rec["dE"][:k] - dE) <= 1e-10 * np.maximum(1.0, np.abs(dE)) + 1e-14 * abs(g("E")[0]))
    if k == n:
        q, ty = r.atoms.identity()
        assert np.array_equal(ty, g("type0_end")) and np.array_equal(q, g("q_end"))
        assert _rel(st["E"], g("E_end")[0]) <= 1e-9


def _phases_case():
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2, n_salt=10)
    return s, p, S.McParams.at_ph(7.0)


P1_KEYS = ("attempted", "accepted", "accepted_forward", "accepted_backward", "blocks_done", "Nh", "Nprot",
           "target_count")


def test_oracle_reproduces_golden_phases():
    z, meta = _load()
    ph = meta["phases"]
    s, p, mc = _phases_case()
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    o.nlist()
    ro = o.phase1_run(mc, ph["seed"], ph["target_pct"], gc_block=ph["gc_block"],
                      epoch_attempt_max=ph["epoch_attempt_max"])
    assert [ro[k] for k in P1_KEYS] == list(z["p1_counts"])
    assert ro["mu_eff"] == z["p1_mu"][0]
    assert np.array_equal(ro["type0"], z["p1_type0"]) and np.array_equal(ro["rng_state"], z["p1_rng"])
    assert np.array_equal(ro["handover"], z["p1_handover"])
    s1 = S.System(s.x, s.y, s.z, ro["q"], ro["type0"], s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)
    o2 = O.Oracle(s1, p, S.bond_adjacency(s.natoms, s.bonds))
    o2.nlist()
    r2 = o2.phase2_run(mc, ro["rng_state"], ph["phase2_trials"], ph["patience"], pools=ro["handover"],
                       pool_sizes=ro["handover_sizes"])
    assert np.array_equal(r2["atoms"], z["p2_atoms"]) and np.array_equal(r2["acc"], z["p2_acc"])
    assert np.all(np.abs(r2["dE"] - z["p2_dE"]) <= 1e-12 * np.maximum(1.0, np.abs(z["p2_dE"])))
    assert np.array_equal(r2["type0"], z["p2_type0"]) and r2["mode"] == z["p2_mode"][0]


@pytest.mark.gpu
def test_device_matches_golden_phases(gpu):
    z, meta = _load()
    ph = meta["phases"]
    s, p, mc = _phases_case()
    r = S.Replica(s, p, mc, seed=ph["seed"], pending_max=8, recompute_interval=400)
    g = r.mc.phase1_run(ph["target_pct"], gc_block=ph["gc_block"], epoch_attempt_max=ph["epoch_attempt_max"])
    assert [g[k] for k in P1_KEYS] == list(z["p1_counts"])
    assert g["mu_eff"] == z["p1_mu"][0]
    q, t = r.atoms.identity()
    assert np.array_equal(t, z["p1_type0"]) and np.array_equal(r.mc.rng_state(), z["p1_rng"])
    n2 = ph["phase2_trials"]
    r.mc.phase2_begin(patience=ph["patience"], record_cap=n2)
    st = r.mc.phase2_run(n2)
    rec = r.mc.phase2_record(st["attempted"])
    assert st["attempted"] == n2 and st["mode"] == z["p2_mode"][0]
    assert np.array_equal(rec["atoms"], z["p2_atoms"]) and np.array_equal(rec["acc"], z["p2_acc"])
    dE = z["p2_dE"]
    assert np.all(np.abs(rec["dE"] - dE) <= 1e-10 * np.maximum(1.0, np.abs(dE)) + 1e-14 * abs(z["p2_E"][0]))
    q, t = r.atoms.identity()
    assert np.array_equal(t, z["p2_type0"])
