"""Pins of the phase-2 restatement (oracle/oracle.cpp or_phase2_run,
phase2_swap_jitter MC/src/mc.cu:159-315).

  - exact enumeration: with the patience out of reach the chain samples the
    canonical distribution exp(-beta E) over the states its moves connect, in
    4-point mode (3 sites / 3 ions, one proton and one Cl to place: 9 states)
    and in W-Cl-only mode (all sites protonated: 3 placements of the W);
  - mode choice and the no-proposal stop;
  - the patience stop: the run ends at the first trial after which `patience`
    trials in a row brought no E < best - best_eps, as recomputed from its own
    record of accepted moves and energies.
"""
import itertools
import math

import numpy as np

from oracle import pyoracle as O
from ph_brush_amd import system as S


def _mt_state(seed):
    x = [seed & 0xFFFFFFFF]
    for i in range(1, 624):
        x.append((1812433253 * (x[-1] ^ (x[-1] >> 30)) + i) & 0xFFFFFFFF)
    return np.array(x + [624], np.uint32)


def _system(site_prot, ion_cl):
    ns, ni = len(site_prot), len(ion_cl)
    rng = np.random.default_rng(11)
    pos = rng.uniform(0.5, 5.5, (ns + ni, 3))
    t = [S.TYPE_NH_P if p else S.TYPE_NH_N for p in site_prot] + [S.TYPE_CL if c else S.TYPE_W for c in ion_cl]
    q = [1.0 if p else 0.0 for p in site_prot] + [-1.0 if c else 0.0 for c in ion_cl]
    s = S.System(pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), np.array(q), np.array(t, np.int32) - 1,
                 np.zeros((0, 2), np.int32), 6.0, 6.0, 6.0, 6.0)
    p = S.Params(rc=2.9, ewald_alpha=0.9, pme_order=4, grid=(16, 16, 16), pme_3dc=False)
    return s, p


def _canonical(s, p, ns, ni, n_prot, n_cl, beta):
    o = O.Oracle(s, p)
    o.nlist()
    states, Es = [], []
    for sites in itertools.combinations(range(ns), n_prot):
        for ions in itertools.combinations(range(ni), n_cl):
            qv = np.array([1.0 if k in sites else 0.0 for k in range(ns)] +
                          [-1.0 if k in ions else 0.0 for k in range(ni)])
            tv = np.array([S.TYPE_NH_P if k in sites else S.TYPE_NH_N for k in range(ns)] +
                          [S.TYPE_CL if k in ions else S.TYPE_W for k in range(ni)], np.int32) - 1
            o.set_charges(qv, tv)
            states.append((sites, ions))
            Es.append(o.total_energy()[0])
    w = np.exp(-beta * (np.array(Es) - min(Es)))
    return states, w / w.sum()


def _frequencies(r, s, ns, states):
    prot = [k for k in range(ns) if s.q[k] > 0]
    cl = [k - ns for k in range(ns, s.natoms) if s.q[k] < 0]
    cur_p, cur_c = set(prot), set(cl)
    counts = np.zeros(len(states))
    for k in range(r["n"]):
        if r["acc"][k]:
            a = r["atoms"][k]
            if r["mode"] == 4:
                cur_p.discard(a[0]); cur_p.add(a[1]); cur_c.discard(a[2] - ns); cur_c.add(a[3] - ns)
            else:
                cur_c.discard(a[0] - ns); cur_c.add(a[1] - ns)
        counts[states.index((tuple(sorted(cur_p)), tuple(sorted(cur_c))))] += 1
    return counts / r["n"]


def test_phase2_four_point_samples_canonical_distribution():
    sp, ic = [1, 0, 0], [1, 0, 0]
    s, p = _system(sp, ic)
    mc = S.McParams(beta=1.0, mu_eff=0.0)
    states, pi = _canonical(s, p, 3, 3, 1, 1, mc.beta)
    o = O.Oracle(s, p)
    o.nlist()
    n = 60000
    r = o.phase2_run(mc, _mt_state(2024), n, patience=10 ** 9)
    assert r["mode"] == 4 and r["stop"] == 0 and r["n"] == n
    assert 0 < r["accepted"] < n
    f = _frequencies(r, s, 3, states)
    assert np.abs(f - pi).max() < 0.015, (f, pi)


def test_phase2_wc_only_samples_canonical_distribution():
    sp, ic = [1, 1], [1, 1, 0]
    s, p = _system(sp, ic)
    mc = S.McParams(beta=1.0, mu_eff=0.0)
    states, pi = _canonical(s, p, 2, 3, 2, 2, mc.beta)
    o = O.Oracle(s, p)
    o.nlist()
    n = 40000
    r = o.phase2_run(mc, _mt_state(7), n, patience=10 ** 9)
    assert r["mode"] == 2 and r["stop"] == 0
    assert np.all(r["atoms"][:, 2:] == -1)
    f = _frequencies(r, s, 2, states)
    assert np.abs(f - pi).max() < 0.015, (f, pi)


def test_phase2_no_valid_proposals():
    s, p = _system([1, 1], [1, 1])          # all sites protonated, no W
    o = O.Oracle(s, p)
    o.nlist()
    r = o.phase2_run(S.McParams(beta=1.0, mu_eff=0.0), _mt_state(1), 100, patience=10)
    assert r["mode"] == 0 and r["stop"] == 2 and r["n"] == 0


def test_phase2_patience_stop_rule():
    sp, ic = [1, 0, 0], [1, 0, 0]
    s, p = _system(sp, ic)
    mc = S.McParams(beta=1.0, mu_eff=0.0)
    o = O.Oracle(s, p)
    o.nlist()
    E0 = o.total_energy()[0]
    for patience in (1, 3, 8, 25):
        r = o.phase2_run(mc, _mt_state(31 + patience), 5000, patience=patience)
        assert r["stop"] == 1
        E, best, since = E0, E0, 0
        for k in range(r["n"]):
            if r["acc"][k]:
                E = E + r["dE"][k]
                if E < best - 1e-9:
                    best, since = E, 0
                else:
                    since += 1
            else:
                since += 1
            assert (since >= patience) == (k == r["n"] - 1), (patience, k)
        assert math.isclose(best, r["best_E"], rel_tol=1e-12, abs_tol=1e-12)
