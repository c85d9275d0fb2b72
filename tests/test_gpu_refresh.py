"""The two potential-grid refreshes agree: the direct one (pending charge
changes convolved with the real-space mesh kernel in one pass over the grid)
and the k-space one (their structure factor times the influence function,
inverse FFT onto the grid) compute the same circular convolution.

Bars, on the same seed:
  - FP64 chain: identical trial sequence (direction, site, ion, accept) and final
    identities; per-move Delta-E within 1e-10 * max(|dE|, 1) + 1e-14 * |E|;
    running energy equal to a full recompute within 1e-9 relative.  Grids whose
    y and z extents are not multiples of the 8 x 8 tile, odd and even spline
    orders, and up to 32 pending changes per refresh are covered;
  - FP32 form (the direct refresh writes its FP32 potential copy in the same
    pass): per-move Delta-E within 1e-4 of the k-space run's over the common
    prefix of the two sequences.
"""
import numpy as np
import pytest

from ph_brush_amd import _lib as L
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu

CASES = [
    # (grid, order, pending_max, refresh of the direct run, trials)
    ((16, 16, 48), 5, 8, L.PHB_REFRESH_AUTO, 3000),
    ((20, 20, 50), 4, 16, L.PHB_REFRESH_DIRECT, 3000),
    ((16, 20, 45), 6, 32, L.PHB_REFRESH_DIRECT, 3000),
    ((20, 24, 54), 3, 4, L.PHB_REFRESH_DIRECT, 3000),
]


def _pair(grid, order, pend, mode, precision=L.PHB_FP64, n=3000):
    s, p, _ = S.config_tiny(order=order, grid=grid, n_perm=2, n_salt=10)
    mc = S.McParams.at_ph(7.9)
    out = []
    for m in (mode, L.PHB_REFRESH_KSPACE):
        out.append(S.Replica(s, p, mc, seed=21, pending_max=pend, recompute_interval=0, record_cap=n,
                             precision=precision, refresh=m))
    return mc, out


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-o{c[1]}-pend{c[2]}" for c in CASES])
def test_direct_refresh_equals_kspace(gpu, case):
    grid, order, pend, mode, n = case
    _, (rd, rk) = _pair(grid, order, pend, mode, n=n)
    sd, sk = rd.mc.run(n), rk.mc.run(n)
    assert sd["refreshes"] > 10 and sd["accepted"] > 50
    gd, gk = rd.mc.record(n), rk.mc.record(n)
    for key in ("fwd", "i", "j", "acc"):
        assert np.array_equal(gd[key], gk[key]), key
    tol = 1e-10 * np.maximum(np.abs(gk["dE"]), 1.0) + 1e-14 * abs(sk["E"])
    assert np.all(np.abs(gd["dE"] - gk["dE"]) <= tol), np.abs(gd["dE"] - gk["dE"]).max()
    qd, td = rd.atoms.identity()
    qk, tk = rk.atoms.identity()
    assert np.array_equal(td, tk) and np.array_equal(qd, qk)
    for r, st in ((rd, sd), (rk, sk)):
        e_re = r.mc.recompute()
        assert abs(st["E"] - e_re) <= 1e-9 * max(1.0, abs(e_re)), (st["E"], e_re)


def test_direct_refresh_fp32_form(gpu):
    n = 5000
    _, (rd, rk) = _pair((16, 16, 48), 5, 8, L.PHB_REFRESH_DIRECT, precision=L.PHB_FP32, n=n)
    rd.mc.run(n)
    rk.mc.run(n)
    gd, gk = rd.mc.record(n), rk.mc.record(n)
    k = 0
    while k < n and all(gd[key][k] == gk[key][k] for key in ("fwd", "i", "j", "acc")):
        k += 1
    assert k > n // 2
    assert np.abs(gd["dE"][:k] - gk["dE"][:k]).max() < 1e-4
    e_run = rd.mc.stats()["E"]
    assert abs(e_run - rd.mc.recompute()) < 1e-2


def test_refresh_tile_shapes_serial(gpu):
    """The direct refresh with 4 x 2, 8 x 2 and 4 x 4-point tiles, switched in the
    middle of a run with changes pending (ragged y and z tiles: a 16 x 20 x 45
    grid): the same trial sequence as the default 8 x 4 tiles, Delta-E and the
    running energy to the bars above (a switch folds the pending changes in at
    once, not at the next refresh); other shapes are refused."""
    s, p, _ = S.config_tiny(order=5, grid=(16, 20, 45), n_perm=2, n_salt=10)
    mc = S.McParams.at_ph(7.9)
    n = 3000
    ref = S.Replica(s, p, mc, seed=21, pending_max=8, recompute_interval=0, record_cap=n,
                    refresh=L.PHB_REFRESH_DIRECT)
    sr = ref.mc.run(n)
    r = S.Replica(s, p, mc, seed=21, pending_max=8, recompute_interval=0, record_cap=n,
                  refresh=L.PHB_REFRESH_DIRECT)
    for ty, tz in ((4, 2), (8, 2), (4, 4)):
        r.mc.set_refresh_tile(ty, tz)
        st = r.mc.run(n // 3)
    assert st["attempted"] == n and st["refreshes"] > 10
    g, gr = r.mc.record(n), ref.mc.record(n)
    for key in ("fwd", "i", "j", "acc"):
        assert np.array_equal(g[key], gr[key]), key
    tol = 1e-10 * np.maximum(np.abs(gr["dE"]), 1.0) + 1e-14 * abs(sr["E"])
    assert np.all(np.abs(g["dE"] - gr["dE"]) <= tol)
    assert abs(st["E"] - sr["E"]) <= 1e-9 * max(1.0, abs(sr["E"]))
    with pytest.raises(L.PhbError):
        r.mc.set_refresh_tile(6, 4)
