"""GPU parity of the phase-2 move rule (phb_phase2_*, phase2_swap_jitter
MC/src/mc.cu:159-315) against the oracle's restatement.

Bar: after a device phase-1 run, phase 2 on the same MC object (pools rebuilt
from the identities, engine stream continued) picks the same atoms and makes the
same accept/reject decisions as the oracle resumed from the device's identities
and engine state; exact Delta-E within 1e-10 * max(|dE|, 1) + 1e-14 |E|; equal
final identities, acceptance count, best energy and stop trial.  Covers the
4-point mode, the W-Cl-only mode (every site protonated) and the patience stop.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu


def _all_protonated(s, mc):
    """Every NH_N site protonated and every W below the pool cut but one turned
    into Cl (the system keeps a net charge of +1, which the net-charge term covers)."""
    q, t = s.q.copy(), s.type0.copy()
    sites = np.flatnonzero(t + 1 == mc.nh_n_type)
    zc = s.z[np.flatnonzero((t + 1 == mc.nh_n_type) | (t + 1 == mc.nh_p_type))].max() + mc.w_z_extra
    ws = [k for k in np.flatnonzero(t + 1 == mc.w_type) if s.z[k] < zc]
    assert len(ws) >= len(sites) >= 2
    for a in sites:
        q[a], t[a] = 1.0, mc.nh_p_type - 1
    for b in ws[:len(sites) - 1]:
        q[b], t[b] = -1.0, mc.cl_type - 1
    return S.System(s.x, s.y, s.z, q, t, s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)


def _run_both(s, p, mc, n1, n2, patience):
    r = S.Replica(s, p, mc, seed=17, pending_max=8, recompute_interval=500, record_cap=max(n1, 1))
    if n1:
        r.mc.run(n1)
    q1, t1 = r.atoms.identity()
    rng = r.mc.rng_state()
    b = r.mc.phase2_begin(patience=patience, record_cap=n2)
    st = r.mc.phase2_run(n2)
    rec = r.mc.phase2_record(st["attempted"])
    s1 = S.System(s.x, s.y, s.z, q1, t1, s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)
    o = O.Oracle(s1, p, r.bonds_host)
    o.nlist()
    ro = o.phase2_run(mc, rng, n2, patience)
    return r, b, st, rec, ro


def _compare(r, st, rec, ro):
    n = st["attempted"]
    assert n == ro["n"] and st["accepted"] == ro["accepted"]
    assert np.array_equal(rec["atoms"], ro["atoms"])
    assert np.array_equal(rec["acc"], ro["acc"])
    tol = 1e-10 * np.maximum(np.abs(ro["dE"]), 1.0) + 1e-14 * abs(ro["E"])
    assert np.all(np.abs(rec["dE"] - ro["dE"]) <= tol), np.abs(rec["dE"] - ro["dE"]).max()
    q, t = r.atoms.identity()
    assert np.array_equal(t, ro["type0"]) and np.array_equal(q, ro["q"])
    assert abs(st["E"] - ro["E"]) <= 1e-9 * max(1.0, abs(ro["E"]))
    assert abs(st["best_E"] - ro["best_E"]) <= 1e-9 * max(1.0, abs(ro["best_E"]))
    e_re = r.mc.recompute()
    assert abs(st["E"] - e_re) <= 1e-9 * max(1.0, abs(e_re))


def test_phase2_four_point_matches_oracle(gpu):
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2, n_salt=10)
    mc = S.McParams.at_ph(7.9)
    r, b, st, rec, ro = _run_both(s, p, mc, 1500, 1500, 10 ** 6)
    assert b["mode"] == 4 and ro["mode"] == 4 and st["stop"] == 0
    assert 0 < st["accepted"] < st["attempted"] == 1500
    _compare(r, st, rec, ro)


def test_phase2_wc_only_matches_oracle(gpu):
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2, n_salt=10)
    mc = S.McParams.at_ph(7.9)
    s = _all_protonated(s, mc)
    r, b, st, rec, ro = _run_both(s, p, mc, 0, 1200, 10 ** 6)
    assert b["mode"] == 2 and ro["mode"] == 2
    assert np.all(rec["atoms"][:, 2:] == -1) and st["accepted"] > 0
    _compare(r, st, rec, ro)


def test_phase2_patience_stop_matches_oracle(gpu):
    s, p, _ = S.config_tiny(order=4, grid=(16, 16, 48), n_perm=2, n_salt=10)
    mc = S.McParams.at_ph(7.0)
    r, b, st, rec, ro = _run_both(s, p, mc, 800, 4000, 40)
    assert st["stop"] == 1 and ro["stop"] == 1 and st["attempted"] < 4000
    _compare(r, st, rec, ro)
    # a further run does nothing once the patience is reached
    st2 = r.mc.phase2_run(100)
    assert st2["attempted"] == st["attempted"]


def test_phase2_records_in_windows(gpu):
    """Records kept window by window (phb_phase2_record_from) equal one record of
    the whole run: the same trials, atoms, decisions and Delta-E."""
    def setup():
        s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2)
        return S.Replica(s, p, S.McParams.at_ph(7.5), seed=9, pending_max=8)
    r1 = setup()
    r1.mc.phase2_begin(patience=10 ** 6, record_cap=700)
    st1 = r1.mc.phase2_run(700)
    whole = r1.mc.phase2_record(st1["attempted"])
    r2 = setup()
    r2.mc.phase2_begin(patience=10 ** 6, record_cap=64)
    parts = []
    done = 0
    while done < 700:
        r2.mc.phase2_record_from(done)
        st2 = r2.mc.phase2_run(min(64, 700 - done))
        n = st2["attempted"] - done
        parts.append(r2.mc.phase2_record(n))
        done = st2["attempted"]
    for key in ("atoms", "acc"):
        assert np.array_equal(np.concatenate([p[key] for p in parts]), whole[key]), key
    np.testing.assert_allclose(np.concatenate([p["dE"] for p in parts]), whole["dE"], rtol=1e-10, atol=1e-10)
