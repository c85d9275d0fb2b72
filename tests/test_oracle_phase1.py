"""Pins of the phase-1 restatement (oracle/oracle.cpp or_phase1_run: build_plan,
apply_direct_protonation and phase_gc_mu_control, MC/src/mc.cu:67-157, :320-546).

  - build_plan's errors: a target below the current protonation, too few W;
  - direct path: exactly the first to_flip NH_N and W (ascending) are flipped, the
    hand-over lists are the flipped and remaining ones, E is the full energy;
  - mu-control path: the run stops at the trial that reaches the target, with the
    ascending hand-over of the final identities; a start at the target returns at
    once with an empty hand-over (the reference fills it only when reached);
  - the initial mu: kT ln(f*/(1-f*)) clamped to +-10 kT (visible when the start is
    the target).
"""
import math

import numpy as np

from oracle import pyoracle as O
from ph_brush_amd import system as S


def _tiny():
    s, p, _ = S.config_tiny(order=4, grid=(16, 16, 48), n_perm=2, n_salt=10)
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    o.nlist()
    return s, p, o


def test_phase1_build_plan_errors():
    s, p, o = _tiny()
    mc = S.McParams.at_ph(7.0)
    t = s.type0 + 1
    cur = np.mean(t[(t == mc.nh_n_type) | (t == mc.nh_p_type)] == mc.nh_p_type) * 100.0
    r = o.phase1_run(mc, 1, target_pct=cur - 10.0)
    assert r["rc"] == -1 and r["err"] == 2
    mc2 = S.McParams(w_z_extra=-1e6)       # no W below the pool cut
    r = o.phase1_run(mc2, 1, target_pct=99.0)
    assert r["rc"] == -1 and r["err"] == 4


def test_phase1_direct_path():
    s, p, o = _tiny()
    mc = S.McParams.at_ph(7.0)
    t = s.type0 + 1
    nh_n = np.flatnonzero(t == mc.nh_n_type)
    nh_p = np.flatnonzero(t == mc.nh_p_type)
    zc = s.z[np.concatenate([nh_n, nh_p])].max() + mc.w_z_extra
    w = np.array([k for k in np.flatnonzero(t == mc.w_type) if s.z[k] < zc])
    nh = len(nh_n) + len(nh_p)
    target = int(math.floor(95.0 / 100.0 * nh))
    k = target - len(nh_p)
    r = o.phase1_run(mc, 1, target_pct=95.0)
    assert r["rc"] == 0 and r["direct"] == 1 and r["attempted"] == k and r["Nprot"] == target
    t1 = r["type0"] + 1
    assert np.all(t1[nh_n[:k]] == mc.nh_p_type) and np.all(t1[nh_n[k:]] == mc.nh_n_type)
    assert np.all(t1[w[:k]] == mc.cl_type) and np.all(t1[w[k:]] == mc.w_type)
    hs = r["handover_sizes"]
    ho = np.split(r["handover"], np.cumsum(hs)[:3])
    assert np.array_equal(ho[0], nh_n[k:]) and np.array_equal(ho[1], nh_n[:k])
    assert np.array_equal(ho[2], w[k:]) and np.array_equal(ho[3], w[:k])
    o.set_charges(r["q"], r["type0"])
    assert abs(o.total_energy()[0] - r["E"]) <= 1e-12 * max(1.0, abs(r["E"]))


def test_phase1_mu_control_reaches_target():
    s, p, o = _tiny()
    mc = S.McParams.at_ph(7.0)
    r = o.phase1_run(mc, 5, target_pct=75.0, gc_block=200, epoch_attempt_max=1000, max_trials=200000)
    assert r["rc"] == 0 and r["direct"] == 0 and r["reached_target"] == 1
    assert r["Nprot"] == r["target_count"] == int(math.floor(0.75 * r["Nh"]))
    t1 = r["type0"] + 1
    hs = r["handover_sizes"]
    ho = np.split(r["handover"], np.cumsum(hs)[:3])
    assert np.array_equal(ho[1], np.flatnonzero(t1 == mc.nh_p_type))
    assert np.array_equal(ho[0], np.flatnonzero(t1 == mc.nh_n_type))
    assert len(ho[1]) == r["target_count"]


def test_phase1_start_at_target():
    s, p, o = _tiny()
    mc = S.McParams.at_ph(7.0)
    t = s.type0 + 1
    nh_p = int(np.sum(t == mc.nh_p_type))
    nh = int(np.sum((t == mc.nh_p_type) | (t == mc.nh_n_type)))
    pct = 100.0 * nh_p / nh + 1e-9
    assert int(math.floor(pct / 100.0 * nh)) == nh_p
    r = o.phase1_run(mc, 1, target_pct=pct)
    assert r["reached_target"] == 1 and r["attempted"] == 0 and r["handover_sizes"].sum() == 0
    fstar = min(max(pct / 100.0, 1e-6), 1 - 1e-6)
    assert r["mu_eff"] == max(-10.0, min(10.0, math.log(fstar / (1.0 - fstar))))


def test_phase1_mu_moves_toward_target():
    """One epoch (Amax = gc_block = max_trials) below the target raises mu, by at
    most kT eta |logit f - logit f*| (the information weight is at most 1); a
    wrong sign or a missing kT factor fails one of the two bounds."""
    s, p, o = _tiny()
    mc = S.McParams.at_ph(7.0)
    pct = 75.0
    r = o.phase1_run(mc, 3, target_pct=pct, gc_block=150, epoch_attempt_max=150, max_trials=150, mu_eta=0.3)
    assert r["rc"] == 0 and r["direct"] == 0 and r["reached_target"] == 0
    assert r["attempted"] == 150 and r["blocks_done"] == 1 and r["accepted"] > 0
    fstar = pct / 100.0
    f = r["Nprot"] / r["Nh"]
    assert f < fstar
    kT = 1.0 / mc.beta
    mu0 = kT * math.log(fstar / (1.0 - fstar))
    bound = kT * 0.3 * abs(math.log(f / (1.0 - f)) - math.log(fstar / (1.0 - fstar)))
    assert 0.0 < r["mu_eff"] - mu0 <= bound * (1 + 1e-12)


def test_phase1_target_on_a_block_end_returns_there():
    """Within one epoch the chain does not depend on gc_block, so with gc_block equal
    to the trial count at which the target is reached, that trial is the block's last:
    the run returns there with no mu update (MC/src/mc.cu:456-468 returns from inside
    the block loop, before the chunk line and the epoch's mu step)."""
    s, p, o = _tiny()
    mc = S.McParams.at_ph(7.0)
    kw = dict(epoch_attempt_max=10 ** 7, nacc_frac=1.0, max_trials=10 ** 6)
    r0 = o.phase1_run(mc, 6, target_pct=75.0, gc_block=10 ** 6, **kw)
    assert r0["rc"] == 0 and r0["reached_target"] == 1 and r0["blocks_done"] == 0
    T = r0["attempted"]
    r1 = o.phase1_run(mc, 6, target_pct=75.0, gc_block=T, **kw)
    assert r1["rc"] == 0 and r1["reached_target"] == 1
    assert (r1["attempted"], r1["accepted"], r1["blocks_done"]) == (T, r0["accepted"], 0)
    assert r1["mu_eff"] == r0["mu_eff"]
    assert np.array_equal(r1["type0"], r0["type0"]) and np.array_equal(r1["rng_state"], r0["rng_state"])
