"""Batched replicas (phb_batch_*): every replica's chain equals its serial run.

Bar: per replica, the trial sequence (direction, site, ion, accept) and the
final identities of a batched run -- uneven system sizes and grids, launches
captured in a hipGraph or issued directly, one persistent launch per sweep,
with or without autotuning in between -- equal those of phb_mc_run on a fresh copy with the same seed and the
same number of trials; the running energy equals a full recompute.
"""
import numpy as np
import pytest

from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu


def _replicas(record_cap):
    specs = [dict(chains=(3, 3), chain_len=6, n_salt=6, grid=(16, 16, 48), ph=7.5, seed=3),
             dict(chains=(4, 3), chain_len=8, n_salt=10, grid=(16, 16, 64), ph=7.0, seed=4),
             dict(chains=(3, 2), chain_len=5, n_salt=2, grid=(16, 16, 48), ph=8.0, seed=5),
             dict(chains=(4, 4), chain_len=7, n_salt=12, grid=(20, 20, 64), ph=7.8, seed=6)]
    reps = []
    for k, sp in enumerate(specs):
        s, p, _ = S.config_tiny(seed=sp["seed"], order=5, grid=sp["grid"], chains=sp["chains"],
                                chain_len=sp["chain_len"], n_perm=1, n_salt=sp["n_salt"],
                                box=(2.0 * sp["chains"][0], 2.0 * sp["chains"][1], 2.5 * sp["chain_len"]))
        mc = S.McParams.at_ph(sp["ph"])
        reps.append((s, p, mc, 100 + k))
    return reps


def _make(reps, record_cap, **kw):
    out = []
    for s, p, mc, seed in reps:
        out.append(S.Replica(s, p, mc, seed=seed, pending_max=8, recompute_interval=0, record_cap=record_cap, **kw))
    return out


def _check_against_serial(reps, batched, record_cap):
    serial = _make(reps, record_cap)
    for rb, rs in zip(batched, serial):
        n = rb.mc.stats()["attempted"]
        assert n > 0
        rs.mc.run(n)
        gb, gs = rb.mc.record(min(n, record_cap)), rs.mc.record(min(n, record_cap))
        for key in ("fwd", "i", "j", "acc"):
            assert np.array_equal(gb[key], gs[key]), key
        qb, tb = rb.atoms.identity()
        qs, ts = rs.atoms.identity()
        assert np.array_equal(tb, ts) and np.array_equal(qb, qs)
        sb, ss = rb.mc.stats(), rs.mc.stats()
        assert sb["accepted"] == ss["accepted"] and sb["nprot"] == ss["nprot"]
        assert abs(sb["E"] - ss["E"]) <= 1e-9 * max(1.0, abs(ss["E"]))
        assert abs(sb["E"] - rb.mc.recompute()) <= 1e-9 * max(1.0, abs(sb["E"]))


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_serial(gpu, use_graph):
    reps = _replicas(4000)
    batched = _make(reps, 4000)
    b = S.Batch([r.mc for r in batched], graph_pairs=3, use_graph=use_graph)
    st = b.run(6)
    assert st["launches"] > 0 and st["attempted"] == sum(r.mc.stats()["attempted"] for r in batched)
    for r in batched:
        assert r.mc.stats()["attempted"] >= 6 * r.mc.stats()["nsites"]
    _check_against_serial(reps, batched, 4000)


def test_persistent_sweep_equals_serial(gpu):
    """The persistent form: one cooperative launch per sweep, each replica's chain
    running to its target and its refreshes folded tile by tile inside the
    launch; a switch to graph replays and back between sweeps leaves every
    sequence exact as well."""
    reps = _replicas(6000)
    batched = _make(reps, 6000)
    b = S.Batch([r.mc for r in batched], graph_pairs=2, persistent=True)
    st = b.run(3)
    assert st["persistent"] == 1 and st["launches"] == 3 and st["replays"] == 3
    b.mode(False)
    st = b.run(2)
    assert st["persistent"] == 0 and st["launches"] > 0
    b.mode(True)
    b.tune(16, 4, 2, 0)
    st = b.run(2)
    assert st["persistent"] == 1 and st["launches"] == 2 and st["pending_max"] == 4
    _check_against_serial(reps, batched, 6000)


@pytest.mark.parametrize("launch_cap", [7, 40])
def test_launch_cap_keeps_every_chain_exact(gpu, launch_cap):
    """A per-launch trial cap (launches end early, mid-budget, some with no
    acceptance) leaves every replica's sequence equal to its serial run."""
    reps = _replicas(6000)
    batched = _make(reps, 6000)
    b = S.Batch([r.mc for r in batched], graph_pairs=2)
    b.tune(8, 8, 2, launch_cap)
    st = b.run(4)
    assert st["launch_cap"] == launch_cap and st["launches"] > 0
    _check_against_serial(reps, batched, 6000)


def test_autotune_keeps_every_chain_exact(gpu):
    reps = _replicas(20000)
    batched = _make(reps, 20000)
    b = S.Batch([r.mc for r in batched])
    best = b.autotune(1)
    print("autotune:", best)
    assert best["spec_window"] in (4, 8, 16) and best["pending_max"] in (2, 4, 8, 16)
    assert best["graph_pairs"] in (2, 4, 8) and best["trials_per_s"] > 0
    assert best["launch_cap"] in (0, 16, 32, 64, 128, 256)
    assert best["refresh_ty"] in (8, 4) and best["refresh_tz"] in (4, 2)
    b.run(2)
    _check_against_serial(reps, batched, 20000)


@pytest.mark.parametrize("persistent", [False, True])
def test_refresh_tile_shapes_keep_every_chain_exact(gpu, persistent):
    """Direct-refresh tiles of 4 x 2, 8 x 2, 4 x 4 and 8 x 4 points, changed
    between sweeps (pending changes folded first), in the graph and the
    persistent forms: every replica equals its serial run at the default tile."""
    reps = _replicas(8000)
    batched = _make(reps, 8000)
    b = S.Batch([r.mc for r in batched], graph_pairs=2, persistent=persistent)
    for ty, tz in ((4, 2), (8, 2), (4, 4), (8, 4)):
        b.set_refresh_tile(ty, tz)
        st = b.run(1)
        assert st["refresh_ty"] == ty and st["refresh_tz"] == tz
        assert st["persistent"] == int(persistent)
    _check_against_serial(reps, batched, 8000)
