"""FP32 form: pair sums, spreading, FFT and convolution in FP32; reductions,
corrections and the Metropolis test in FP64.  Errors against FP64 are stated
and bounded here:

  - total energy (FP32 PME plan + FP32 pair sums) vs the FP64 oracle:
    relative error below 2e-5 on the 2,640-atom brush;
  - per-move Delta-E of the FP32 chain vs the FP64 chain (itself equal to the
    oracle within 1e-10) over 20,000 sampled moves: RMS below 2e-4 kT, max below
    2e-3 kT; 100 of them also directly against the oracle;
  - the FP32 chain's decisions equal the FP64 chain's except at a trial whose
    Metropolis exponent is within the FP32 Delta-E error of its threshold.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import _lib as L
from ph_brush_amd import surrogate as SG
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu


def test_total_energy_fp32_form(gpu):
    s, p, _ = S.config_small_fp64()
    atoms = S.DeviceAtoms(s)
    bonds = S.DeviceCSR.from_host(*S.bond_adjacency(s.natoms, s.bonds))
    box = S.box_of(s)
    neigh = S.build_half_neighbors_gpu(atoms, box, bonds, p.rc)
    params = S.DeviceParams(p)
    pme32 = S.PMEPlan(params, box, p.rc, precision=L.PHB_FP32)
    e32, _ = S.total_energy(atoms, params, neigh, pme32, box)
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    o.nlist()
    e64, _ = o.total_energy()
    rel = abs(e32 - e64) / abs(e64)
    print(f"total energy FP32 form: {e32!r} vs FP64 oracle {e64!r}, relative error {rel:.2e}")
    assert rel < 2e-5


def _pair(seed_ph=7.5):
    s, p, _ = S.config_small_fp64()
    mc = S.McParams.at_ph(seed_ph)
    r64 = S.Replica(s, p, mc, seed=5, pending_max=8, recompute_interval=0, record_cap=20000)
    r32 = S.Replica(s, p, mc, seed=5, pending_max=8, recompute_interval=0, record_cap=20000, precision=L.PHB_FP32)
    return s, p, mc, r64, r32


def test_delta_e_fp32_form_error(gpu):
    s, p, mc, r64, r32 = _pair()
    for r in (r64, r32):
        r.mc.run(3000)
    q64, t64 = r64.atoms.identity()
    q32, t32 = r32.atoms.identity()
    assert np.array_equal(t64, t32)
    rng = np.random.default_rng(8)
    i, j, f = SG.sample_moves(rng, t64, s.z, mc, 20000)
    d64 = r64.mc.delta_e(i, j, f)
    d32 = r32.mc.delta_e(i, j, f)
    err = d32 - d64
    rms, mx = float(np.sqrt(np.mean(err ** 2))), float(np.abs(err).max())
    print(f"Delta-E FP32 form vs FP64 over 20000 moves: rms {rms:.2e} kT, max {mx:.2e} kT "
          f"(|dE| rms {np.sqrt(np.mean(d64 ** 2)):.2f} kT)")
    assert rms < 2e-4 and mx < 2e-3
    # 100 of them directly against the oracle's full recompute
    o = O.Oracle(s, p, r64.bonds_host)
    o.nlist()
    o.set_charges(q64, t64)
    e0, _ = o.total_energy()
    for k in range(100):
        qq, tt = q64.copy(), t64.copy()
        if f[k]:
            qq[i[k]], tt[i[k]], qq[j[k]], tt[j[k]] = 1.0, mc.nh_p_type - 1, -1.0, mc.cl_type - 1
        else:
            qq[i[k]], tt[i[k]], qq[j[k]], tt[j[k]] = 0.0, mc.nh_n_type - 1, 0.0, mc.w_type - 1
        o.set_charges(qq, tt)
        e1, _ = o.total_energy()
        assert abs(d32[k] - (e1 - e0)) < 2e-3
        assert abs(d64[k] - (e1 - e0)) <= 1e-10 * max(1.0, abs(e1 - e0)) + 1e-14 * abs(e0)
    o.set_charges(q64, t64)


def test_chain_fp32_form_decisions(gpu):
    s, p, mc, r64, r32 = _pair(7.9)
    n = 20000
    r64.mc.run(n)
    r32.mc.run(n)
    g64, g32 = r64.mc.record(n), r32.mc.record(n)
    k = 0
    while k < n and all(g64[key][k] == g32[key][k] for key in ("fwd", "i", "j", "acc")):
        k += 1
    if k < n:
        # a split may only happen where the FP64 decision was within the FP32 error of its threshold
        ex = -(g64["dE"][k] - (1 if g64["fwd"][k] else -1) * mc.mu_eff)   # beta = 1; logH differs little
        u = g64["u"][k]
        thr = 0.0 if math.isnan(u) else math.log(u)
        assert abs(abs(g64["dE"][k] - g32["dE"][k])) < 2e-3
        print(f"FP32/FP64 chains split at trial {k} (margin to threshold ~ {abs(ex - thr):.2e})")
    else:
        q64, t64 = r64.atoms.identity()
        q32, t32 = r32.atoms.identity()
        assert np.array_equal(t64, t32)
    m = min(k, n)
    assert np.abs(g64["dE"][:m] - g32["dE"][:m]).max() < 2e-3
    e_run = r32.mc.stats()["E"]
    e_re = r32.mc.recompute()
    assert abs(e_run - e_re) < 1e-2, (e_run, e_re)
