"""GPU parity of the device MC chain against the oracle's reference algorithm
(full energy recompute per trial, std::mt19937 + libstdc++ distributions).

Bars (FP64 plain path):
  - per-move Delta-E:  |dE_gpu - dE_oracle| <= 1e-10 * max(|dE|, 1) + 1e-14 * |E_total|
    (the second term is the rounding floor of the oracle's difference of two totals);
  - accept/reject sequence, directions and picked atoms: identical, except that a
    divergence is allowed only at a move whose Metropolis exponent lies within
    1e-9 of its threshold (0, or ln u) -- after such a move the streams differ;
  - final identities identical; running energy equal to a full recompute.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import _lib as L
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu


def _tol(dE, Etot):
    return 1e-10 * max(abs(dE), 1.0) + 1e-14 * abs(Etot)


def _valid_moves(s, mc, n, seed):
    rng = np.random.default_rng(seed)
    t = s.type0 + 1
    nhn, nhp = np.where(t == mc.nh_n_type)[0], np.where(t == mc.nh_p_type)[0]
    w, cl = np.where(t == mc.w_type)[0], np.where(t == mc.cl_type)[0]
    mv = []
    for k in range(n):
        if (k % 2 == 0 and len(nhn) and len(w)) or not (len(nhp) and len(cl)):
            mv.append((int(rng.choice(nhn)), int(rng.choice(w)), 1))
        else:
            mv.append((int(rng.choice(nhp)), int(rng.choice(cl)), 0))
    return mv


def _oracle_delta(o, s, mc, i, j, f):
    e0, _ = o.total_energy()
    q, t = s.q.copy(), s.type0.copy()
    if f:
        q[i], t[i], q[j], t[j] = 1.0, mc.nh_p_type - 1, -1.0, mc.cl_type - 1
    else:
        q[i], t[i], q[j], t[j] = 0.0, mc.nh_n_type - 1, 0.0, mc.w_type - 1
    o.set_charges(q, t)
    e1, _ = o.total_energy()
    o.set_charges(s.q, s.type0)
    return e1 - e0, e0


@pytest.mark.parametrize("name", ["tiny_o4", "tiny_o5_lj", "small_2640"])
def test_delta_e_matches_full_recompute(gpu, name):
    if name == "tiny_o4":
        s, p, mc = S.config_tiny(order=4, grid=(16, 16, 48))
    elif name == "tiny_o5_lj":
        s, p, mc = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2)
        sig = np.ones((8, 8)); sig[S.TYPE_CL - 1, :] = sig[:, S.TYPE_CL - 1] = 1.15
        sig[S.TYPE_NH_P - 1, :] = sig[:, S.TYPE_NH_P - 1] = 0.9
        p.lj_sig = sig.ravel()
        p.lj_rc = 1.6    # LJ beyond WCA so that identity changes alter LJ
    else:
        s, p, mc = S.config_small_fp64()
    r = S.Replica(s, p, mc, seed=5, pending_max=16)
    o = O.Oracle(s, p, r.bonds_host)
    o.nlist()
    moves = _valid_moves(s, mc, 24, 3)
    got = r.mc.delta_e([m[0] for m in moves], [m[1] for m in moves], [m[2] for m in moves])
    for (i, j, f), g in zip(moves, got):
        exp, e0 = _oracle_delta(o, s, mc, i, j, f)
        assert abs(g - exp) <= _tol(exp, e0), (i, j, f, g, exp)


def _compare_chains(rec_g, rec_o, n):
    """Index of the first divergence (n if none); asserts it is near threshold."""
    for k in range(n):
        same = (rec_g["fwd"][k] == rec_o["fwd"][k] and rec_g["i"][k] == rec_o["i"][k]
                and rec_g["j"][k] == rec_o["j"][k] and rec_g["acc"][k] == rec_o["acc"][k])
        if not same:
            assert rec_g["fwd"][k] == rec_o["fwd"][k] and rec_g["i"][k] == rec_o["i"][k] \
                and rec_g["j"][k] == rec_o["j"][k], f"proposal diverged at trial {k} without a prior decision split"
            ex = rec_o["expo"][k]
            u = rec_o["u"][k]
            thr = 0.0 if math.isnan(u) else math.log(u)
            assert abs(ex - thr) < 1e-9 or abs(ex) < 1e-9, f"decision split at trial {k} far from threshold"
            return k
    return n


@pytest.mark.parametrize("case", [("tiny", 3000, 4), ("tiny", 3000, 32), ("small", 1200, 32), ("tiny_coul_cut", 3000, 8)])
def test_chain_matches_reference_sequence(gpu, case):
    """tiny_coul_cut: total_energy's coul_cut branch (MC/src/mc.cu:52-56) -- LJ plus
    elF q q / r over the listed pairs, no Ewald parts and no slab term although the
    box is extended -- on the chain's own alpha-0 plan."""
    name, n, pend = case
    if name.startswith("tiny"):
        s, p, mc = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2)
        mc = S.McParams.at_ph(7.9)
        if name == "tiny_coul_cut":
            p.coulombtype = L.PHB_COUL_CUT
    else:
        s, p, mc = S.config_small_fp64()
    seed = 1
    r = S.Replica(s, p, mc, seed=seed, pending_max=pend, recompute_interval=250, record_cap=n)
    st = r.mc.run(n)
    assert st["attempted"] == n and st["status"] == 0
    rg = r.mc.record(n)
    o = O.Oracle(s, p, r.bonds_host)
    o.nlist()
    ro = o.gc_run(mc, n, seed)
    assert ro["n"] == n
    k = _compare_chains(rg, ro, n)
    # Delta-E agreement over the common prefix
    for t in range(k):
        assert abs(rg["dE"][t] - ro["dE"][t]) <= _tol(ro["dE"][t], ro["E"]), t
    if k == n:
        q, ty = r.atoms.identity()
        assert np.array_equal(q, ro["q"]) and np.array_equal(ty, ro["type0"])
        assert abs(st["E"] - ro["E"]) <= 1e-9 * max(1.0, abs(ro["E"]))
    # accepted moves exist in both directions (the chain is exercised)
    assert st["accepted_forward"] > 0 and st["accepted_backward"] > 0
    # running energy vs a fresh recompute
    e_run = r.mc.stats()["E"]
    e_re = r.mc.recompute()
    assert abs(e_run - e_re) <= 1e-9 * max(1.0, abs(e_re))
    e_tot, _ = r.total_energy()
    assert abs(e_tot - e_re) <= 1e-10 * max(1.0, abs(e_re))


def test_chain_matches_reference_sequence_44k(gpu):
    """The chain the benchmark times, at its size: the 44,000-atom diblock brush
    (SPME 64x64x256, order 5, pH 7), pending budget 8 and window 8 as in the
    benchmark, 400 trials against the oracle's full-recompute chain (the same
    bar as above; several refreshes of the potential grid happen in between)."""
    s, p, mc = S.config_30k_fp64()
    n, seed = 400, 1
    r = S.Replica(s, p, mc, seed=seed, pending_max=8, spec_window=8, recompute_interval=0, record_cap=n)
    st = r.mc.run(n)
    assert st["attempted"] == n and st["status"] == 0 and st["accepted"] >= 4
    rg = r.mc.record(n)
    o = O.Oracle(s, p, r.bonds_host)
    o.nlist()
    ro = o.gc_run(mc, n, seed)
    k = _compare_chains(rg, ro, n)
    for t in range(k):
        assert abs(rg["dE"][t] - ro["dE"][t]) <= _tol(ro["dE"][t], ro["E"]), t
    if k == n:
        q, ty = r.atoms.identity()
        assert np.array_equal(q, ro["q"]) and np.array_equal(ty, ro["type0"])
        assert abs(st["E"] - ro["E"]) <= 1e-9 * max(1.0, abs(ro["E"]))
    e_re = r.mc.recompute()
    assert abs(st["E"] - e_re) <= 1e-9 * max(1.0, abs(e_re))


def test_chain_bit_reproducible_44k(gpu):
    """Two chains from the same state and seed on the 44,000-atom brush, 6,000
    trials with a drift reset every 1,000 (each rebuilds the potential grid from
    the spread): the same trial records, identities and energies bit for bit."""
    s, p, mc = S.config_30k_fp64()
    n = 6000
    out = []
    for _ in range(2):
        r = S.Replica(s, p, mc, seed=5, pending_max=8, spec_window=8, recompute_interval=1000, record_cap=n)
        st = r.mc.run(n)
        assert st["attempted"] == n and st["status"] == 0 and st["recomputes"] >= 5
        out.append((st, r.mc.record(n), r.atoms.identity()))
        del r
    (s1, r1, (q1, t1)), (s2, r2, (q2, t2)) = out
    for k in ("acc", "i", "j", "fwd"):
        assert np.array_equal(r1[k], r2[k]), k
    assert np.array_equal(r1["dE"].view(np.int64), r2["dE"].view(np.int64))
    assert np.array_equal(q1, q2) and np.array_equal(t1, t2)
    for k in ("E", "E_recomputed", "max_drift"):
        assert np.float64(s1[k]).view(np.int64) == np.float64(s2[k]).view(np.int64), k


def test_set_mu_equals_a_chain_built_at_that_mu(gpu):
    """A chain moved to another pH by phb_mc_set_mu runs the trial sequence of a
    chain created at that pH (same state and seed)."""
    s, p, _ = S.config_small_fp64()
    n = 2000
    mc8 = S.McParams.at_ph(8.0)
    a = S.Replica(s, p, S.McParams.at_ph(6.0), seed=9, record_cap=n)
    a.mc.set_mu(mc8.mu_eff)
    b = S.Replica(s, p, mc8, seed=9, record_cap=n)
    sa, sb = a.mc.run(n), b.mc.run(n)
    ra, rb = a.mc.record(n), b.mc.record(n)
    for k in ("acc", "i", "j", "fwd"):
        assert np.array_equal(ra[k], rb[k]), k
    assert sa["E"] == sb["E"] and sa["nprot"] == sb["nprot"]
    with pytest.raises(L.PhbError):
        a.mc.set_mu(float("nan"))


def test_rng_state_matches_std_mt19937(gpu):
    """The device engine starts from std::mt19937(seed): the C++ standard's seeding
    x[0] = seed, x[i] = 1812433253 * (x[i-1] ^ (x[i-1] >> 30)) + i, index 624."""
    s, p, mc = S.config_tiny()
    r = S.Replica(s, p, mc, seed=42)
    st = r.mc.rng_state()
    x = [42]
    for i in range(1, 624):
        x.append((1812433253 * (x[-1] ^ (x[-1] >> 30)) + i) & 0xFFFFFFFF)
    assert np.array_equal(st[:624], np.array(x, np.uint32)) and st[624] == 624


@pytest.mark.parametrize("form", ["plain", "da"])
def test_running_energy_over_uneven_runs(gpu, form):
    """Runs of uneven length with no intermediate full recompute: launches queued
    beyond a run's target must leave the potential grid untouched, so the
    running energy still equals a fresh recompute after many runs."""
    s, p, _ = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2)
    mc = S.McParams.at_ph(7.9)
    r = S.Replica(s, p, mc, seed=11, pending_max=8, recompute_interval=0)
    if form == "da":
        r.mc.da_attach(r_desc=6.0)
        sur = S.Surrogate(0)
        run = lambda n: r.mc.da_run(sur, n)
    else:
        run = r.mc.run
    for n in (123, 77, 301, 5, 250, 999):
        run(n)
    st = r.mc.stats()
    assert st["attempted"] == 1755 and st["accepted"] > 20
    e_run = st["E"]
    e_re = r.mc.recompute()
    assert abs(e_run - e_re) <= 1e-9 * max(1.0, abs(e_re)), (e_run, e_re)
