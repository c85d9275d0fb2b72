"""GPU parity of phase 1 as main() runs it (phb_phase1_run: build_plan checks,
apply_direct_protonation or phase_gc_mu_control on the device chain,
MC/src/mc.cu:67-157, :320-546) against the oracle's restatement.

Bar: identical attempted / accepted / forward / backward counts, epochs, final
protonation and identities, bit-identical mu (the same FP64 update from the same
counts) and engine state afterwards, energy within 1e-9 relative; phase 2 from
the hand-over pools then matches the oracle trial by trial (atoms, decisions,
Delta-E within 1e-10).  Both the mu-control and the direct path.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu


def _both(target_pct, seed=5, **kw):
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2, n_salt=10)
    mc = S.McParams.at_ph(7.0)
    r = S.Replica(s, p, mc, seed=seed, pending_max=8, recompute_interval=400)
    g = r.mc.phase1_run(target_pct, **kw)
    o = O.Oracle(s, p, r.bonds_host)
    o.nlist()
    ro = o.phase1_run(mc, seed, target_pct, **kw)
    return s, p, mc, r, g, o, ro


def _check_phase1(r, g, ro):
    assert ro["rc"] == 0
    for k in ("attempted", "accepted", "accepted_forward", "accepted_backward", "blocks_done", "Nh", "Nprot",
              "target_count", "reached_target", "direct"):
        assert g[k] == ro[k], (k, g[k], ro[k])
    assert g["mu_eff"] == ro["mu_eff"]
    q, t = r.atoms.identity()
    assert np.array_equal(t, ro["type0"]) and np.array_equal(q, ro["q"])
    assert np.array_equal(r.mc.rng_state(), ro["rng_state"])
    assert abs(g["E"] - ro["E"]) <= 1e-9 * max(1.0, abs(ro["E"]))


def _check_phase2(s, p, mc, r, ro, n2=600, patience=10 ** 6):
    r.mc.phase2_begin(patience=patience, record_cap=n2)
    st = r.mc.phase2_run(n2)
    rec = r.mc.phase2_record(st["attempted"])
    s1 = S.System(s.x, s.y, s.z, ro["q"], ro["type0"], s.bonds, s.Lx, s.Ly, s.Lz, s.Lz_data)
    o2 = O.Oracle(s1, p, r.bonds_host)
    o2.nlist()
    r2 = o2.phase2_run(mc, ro["rng_state"], n2, patience, pools=ro["handover"], pool_sizes=ro["handover_sizes"])
    assert st["mode"] == r2["mode"] and st["attempted"] == r2["n"] and st["accepted"] == r2["accepted"]
    assert np.array_equal(rec["atoms"], r2["atoms"]) and np.array_equal(rec["acc"], r2["acc"])
    tol = 1e-10 * np.maximum(np.abs(r2["dE"]), 1.0) + 1e-14 * abs(r2["E"])
    assert np.all(np.abs(rec["dE"] - r2["dE"]) <= tol)
    return st


def _check_events(r, g, ro, mc):
    """The event log behind the report's [sgmc] lines against the oracle's run."""
    from ph_brush_amd import _lib as L
    ev = r.mc.phase1_events()
    ups = [e for e in ev if e["kind"] == L.PHB_P1_MU_UPDATE]
    assert len(ups) == ro["blocks_done"]
    assert [e["epoch"] for e in ups] == list(range(len(ups)))
    last = ev[-1]
    assert last["kind"] == L.PHB_P1_REACHED and last["epoch"] == ro["blocks_done"]
    assert last["nprot"] == ro["Nprot"] and last["remaining"] == 0
    assert (last["attempted"], last["accepted"]) == (ro["attempted"], ro["accepted"])
    assert (last["fwd_total"], last["bwd_total"]) == (ro["accepted_forward"], ro["accepted_backward"])
    assert last["mu_over_kT"] == mc.beta * ro["mu_eff"]           # the same FP64 updates, bit for bit
    blocks = [e for e in ev if e["kind"] in (L.PHB_P1_CHUNK, L.PHB_P1_REACHED)]
    assert sum(e["block_attempted"] for e in blocks) == ro["attempted"]
    assert sum(e["block_accepted"] for e in blocks) == ro["accepted"]
    assert last["step_in_block"] == last["block_attempted"] <= last["gc_block"]


def test_phase1_mu_control_then_phase2(gpu):
    s, p, mc, r, g, o, ro = _both(75.0, gc_block=200, epoch_attempt_max=1000, max_trials=200000)
    assert g["reached_target"] == 1 and g["direct"] == 0 and g["blocks_done"] >= 1
    _check_phase1(r, g, ro)
    _check_events(r, g, ro, mc)
    st = _check_phase2(s, p, mc, r, ro)
    assert st["mode"] == 4 and st["accepted"] > 0


def test_phase1_direct_then_phase2(gpu):
    from ph_brush_amd import _lib as L
    s, p, mc, r, g, o, ro = _both(95.0)
    assert g["direct"] == 1
    _check_phase1(r, g, ro)
    ev = r.mc.phase1_events()
    assert len(ev) == 1 and ev[0]["kind"] == L.PHB_P1_DIRECT and ev[0]["attempted"] == ro["attempted"] > 0
    e_re = r.mc.recompute()
    assert abs(g["E"] - e_re) <= 1e-9 * max(1.0, abs(e_re))
    _check_phase2(s, p, mc, r, ro)


@pytest.mark.parametrize("seed,gc_block,amax,nacc_frac,epochs", [
    (6, 2211, 10 ** 7, 1.0, 0),   # the target lands on the last trial of the first block
    (6, 36, 300, 0.10, 3),        # ... of a block after three mu updates
])
def test_phase1_target_on_last_trial_of_a_block(gpu, seed, gc_block, amax, nacc_frac, epochs):
    """The trial that reaches the target is the last of its block: the run ends
    there (step_in_block = gc_block), with no chunk line and no mu update after it
    (MC/src/mc.cu:456-468 returns from inside the block loop)."""
    from ph_brush_amd import _lib as L
    s, p, mc, r, g, o, ro = _both(75.0, seed=seed, gc_block=gc_block, epoch_attempt_max=amax,
                                  nacc_frac=nacc_frac, max_trials=200000)
    assert ro["reached_target"] == 1 and ro["attempted"] % gc_block == 0 and ro["blocks_done"] == epochs
    _check_phase1(r, g, ro)
    _check_events(r, g, ro, mc)
    ev = r.mc.phase1_events()
    assert ev[-1]["step_in_block"] == gc_block
    assert sum(e["kind"] == L.PHB_P1_REACHED for e in ev) == 1
