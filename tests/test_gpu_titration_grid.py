"""The batched titration grid at its real sizes: 32 replicas (8 pH x 4 salt
levels, seed 7), 5,000-35,000 atoms each, SPME grids 32-64 x 32-64 x 128-256,
advanced together by phb_batch_run with the launch shape the autotuner picks on
this grid (window 16, pending budget 4, 8 launch pairs per graph, launch cap 32).
A quarter of the replicas refresh their potential grid in k-space, the rest
directly; in the persistent form (one cooperative launch per sweep, which needs
the direct refresh) every replica refreshes directly.

Bars: every replica's trial sequence (direction, site, ion, decision), final
identities, counts and running energy equal those of phb_mc_run on a separate
copy with the same seed and trial count; the running energy equals a full
recompute.  Against the oracle (the reference algorithm, MC/src/neighbor.cu:74-125,
MC/src/mc.cu:43-60, :379-503), on the smallest and largest replica and one of
each remaining grid shape: the half neighbour list is bit-exact, the total
energy with the final identities agrees within 1e-10 relative, and the smallest
replica's first 300 batched trials equal the oracle's chain."""
import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import _lib as L
from ph_brush_amd import system as S

from test_gpu_mc import _compare_chains, _tol

pytestmark = pytest.mark.gpu

REC = 40000


@pytest.fixture(scope="module")
def grid():
    g = S.config_titration_grid()
    assert len(g) == 32
    return g


def _make(grid, persistent=False):
    reps = []
    for k, (s, p, mc, _) in enumerate(grid):
        refresh = L.PHB_REFRESH_KSPACE if k % 4 == 1 and not persistent else L.PHB_REFRESH_AUTO
        reps.append(S.Replica(s, p, mc, seed=1000 + k, pending_max=8, recompute_interval=0, record_cap=REC,
                              refresh=refresh))
    return reps


@pytest.mark.parametrize("persistent", [False, True])
def test_titration_grid_batch_equals_serial_and_oracle(gpu, grid, persistent):
    natoms = [g[0].natoms for g in grid]
    assert min(natoms) >= 5000 and max(natoms) <= 35000
    assert len({g[1].grid for g in grid}) >= 6     # 32 and 64 laterally, 128 and 256 along z
    batched = _make(grid, persistent)
    b = S.Batch([r.mc for r in batched], persistent=persistent)
    b.tune(16, 4, 8, 32)
    st = b.run(2)
    assert st["launches"] > 0 and st["replays"] > 0
    if persistent:
        assert st["persistent"] == 1 and st["launches"] == 2
    serial = _make(grid, persistent)
    for k, (rb, rs) in enumerate(zip(batched, serial)):
        sb = rb.mc.stats()
        n = sb["attempted"]
        assert n >= 2 * sb["nsites"] and n <= REC, (k, n)
        rs.mc.run(n)
        gb, gs = rb.mc.record(n), rs.mc.record(n)
        for key in ("fwd", "i", "j", "acc"):
            assert np.array_equal(gb[key], gs[key]), (k, key)
        qb, tb = rb.atoms.identity()
        qs, ts = rs.atoms.identity()
        assert np.array_equal(tb, ts) and np.array_equal(qb, qs), k
        ss = rs.mc.stats()
        assert sb["accepted"] == ss["accepted"] > 0 and sb["nprot"] == ss["nprot"], k
        assert abs(sb["E"] - ss["E"]) <= 1e-9 * max(1.0, abs(ss["E"])), k
        assert abs(sb["E"] - rb.mc.recompute()) <= 1e-9 * max(1.0, abs(sb["E"])), k
    del serial
    # the oracle on the extremes and on one replica of each other grid shape
    pick = {int(np.argmin(natoms)), int(np.argmax(natoms))}
    seen = {grid[k][1].grid for k in pick}
    for k, g in enumerate(grid):
        if g[1].grid not in seen:
            pick.add(k)
            seen.add(g[1].grid)
    for k in sorted(pick):
        s, p, mc, label = grid[k]
        r = batched[k]
        o = O.Oracle(s, p, r.bonds_host)
        h, l = r.neigh.download()
        ho, lo = o.nlist()
        assert np.array_equal(h, ho) and np.array_equal(l, lo), label
        q, t = r.atoms.identity()
        o.set_charges(q, t)
        eo, _ = o.total_energy()
        e, _ = r.total_energy()
        assert abs(e - eo) <= 1e-10 * max(1.0, abs(eo)), (label, e, eo)
        assert abs(r.mc.stats()["E"] - eo) <= 1e-9 * max(1.0, abs(eo)), label
    # the smallest replica's batched chain against the reference algorithm
    k = int(np.argmin(natoms))
    s, p, mc, _ = grid[k]
    o = O.Oracle(s, p, batched[k].bonds_host)
    o.nlist()
    n = 300
    ro = o.gc_run(mc, n, 1000 + k)
    rg = batched[k].mc.record(n)
    m = _compare_chains(rg, ro, n)
    assert m == n
    for t in range(n):
        assert abs(rg["dE"][t] - ro["dE"][t]) <= _tol(ro["dE"][t], ro["E"]), t
