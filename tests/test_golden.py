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
