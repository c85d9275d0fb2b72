"""Pins of the delayed-acceptance oracle (oracle/oracle.cpp or_da_*, oracle/mlp.py).

  or_da_descriptors   numpy brute force over all atoms; move/reverse-move invariance
  or_da_run           exact enumeration of a 6-state titration with a deliberately
                      poor surrogate (stationarity); the proposal stream rebuilt
                      from raw mt19937 output when stage 1 rejects everything
  mlp.bf16            torch's float32 -> bfloat16 conversion (round to nearest even)
  mlp.forward         torch float64 autograd forward
  mlp.loss_and_grad   central finite differences (FP64) and torch autograd
  mlp.adam_step       torch.optim.Adam
"""
import math

import numpy as np
import pytest
import torch

from oracle import mlp as MLP
from oracle import pyoracle as O
from ph_brush_amd import system as S


def _min_image(d, L):
    return d - L * np.rint(d / L)


def _brute_desc(s, species, r_desc, i, j):
    th = O.desc_thresholds(r_desc)
    out = np.zeros(48, np.int64)
    for ctr, sg in ((i, 1), (j, -1)):
        dx = _min_image(s.x[ctr] - s.x, s.Lx)
        dy = _min_image(s.y[ctr] - s.y, s.Ly)
        dz = _min_image(s.z[ctr] - s.z, s.Lz)
        r2 = dx * dx + dy * dy + dz * dz
        sp = species[s.type0]
        ok = (r2 <= r_desc * r_desc) & (sp >= 0)
        ok[[i, j]] = False
        sh = np.searchsorted(th, r2[ok], side="right") - 1
        np.add.at(out, sh * 4 + sp[ok], sg)
    return out


def test_desc_thresholds():
    th = O.desc_thresholds(6.0)
    assert np.array_equal(th, np.array([(k * 6.0 / 12.0) ** 2 for k in range(12)]))


def test_da_descriptors_brute_force_and_reverse_invariance():
    s, p, mc = S.config_tiny(order=4, grid=(16, 16, 48))
    species = S.default_species(p.ntypes)
    o = O.Oracle(s, p)
    rng = np.random.default_rng(3)
    sites = np.flatnonzero((s.type0 == S.TYPE_NH_N - 1) | (s.type0 == S.TYPE_NH_P - 1))
    ions = np.flatnonzero((s.type0 == S.TYPE_W - 1) | (s.type0 == S.TYPE_CL - 1))
    i = rng.choice(sites, 40)
    j = rng.choice(ions, 40)
    d, rij = o.da_descriptors(species, 6.0, i, j)
    for m in range(40):
        assert np.array_equal(d[m], _brute_desc(s, species, 6.0, i[m], j[m]))
        assert abs(rij[m] - math.sqrt(((_min_image(s.x[i[m]] - s.x[j[m]], s.Lx)) ** 2 +
                                       (_min_image(s.y[i[m]] - s.y[j[m]], s.Ly)) ** 2 +
                                       (_min_image(s.z[i[m]] - s.z[j[m]], s.Lz)) ** 2))) < 1e-12
    # a move changes only i and j, which both centres exclude: same descriptor after it
    t2 = s.type0.copy()
    for m in range(40):
        t2[i[m]] = S.TYPE_NH_P - 1 if s.type0[i[m]] == S.TYPE_NH_N - 1 else S.TYPE_NH_N - 1
    t3 = s.type0.copy()
    t3[i[0]] = t2[i[0]]
    t3[j[0]] = S.TYPE_CL - 1 if s.type0[j[0]] == S.TYPE_W - 1 else S.TYPE_W - 1
    o.set_charges(s.q, t3)
    d3, _ = o.da_descriptors(species, 6.0, i[:1], j[:1])
    assert np.array_equal(d3[0], d[0])
    assert d[:, :].any()


def _tiny4():
    x = np.array([1.0, 2.6, 1.2, 3.0]); y = np.array([1.0, 1.3, 2.5, 2.4]); z = np.array([1.0, 1.5, 1.0, 2.2])
    t0 = np.array([S.TYPE_NH_N, S.TYPE_NH_N, S.TYPE_W, S.TYPE_W], np.int32) - 1
    s = S.System(x, y, z, np.zeros(4), t0, np.zeros((0, 2), np.int32), 6.0, 6.0, 6.0, 6.0)
    p = S.Params(rc=2.9, ewald_alpha=0.9, pme_order=4, grid=(16, 16, 16), pme_3dc=False)
    return s, p, t0


@pytest.mark.parametrize("sym", [False, True])
def test_da_chain_samples_exact_distribution(sym):
    """Stage 2 restores the exact distribution whatever the surrogate: here a
    random table that is badly wrong for most moves, antisymmetric under the
    reverse move or not (the general second-stage rule)."""
    s, p, t0 = _tiny4()
    mc = S.McParams(beta=1.0, mu_eff=0.3)
    states = [(a, b, c, d) for a in (0, 1) for b in (0, 1) for c in (0, 1) for d in (0, 1) if a + b == c + d]
    o = O.Oracle(s, p)
    o.nlist()
    Es = []
    for st in states:
        qv = np.array([st[0], st[1], -st[2], -st[3]], float)
        tv = np.array([S.TYPE_NH_P if st[0] else S.TYPE_NH_N, S.TYPE_NH_P if st[1] else S.TYPE_NH_N,
                       S.TYPE_CL if st[2] else S.TYPE_W, S.TYPE_CL if st[3] else S.TYPE_W], np.int32) - 1
        o.set_charges(qv, tv)
        Es.append(o.total_energy()[0])
    o.set_charges(np.zeros(4), t0)
    w = np.array([math.exp(-(E - mc.mu_eff * (st[0] + st[1]))) for E, st in zip(Es, states)])
    pi = w / w.sum()
    rng = np.random.default_rng(5)
    F = rng.normal(0.0, 3.0, (4, 4))
    B = rng.normal(0.0, 2.0, (4, 4)) if sym else None    # a part that does not flip with the move
    n = 80000
    r = o.da_run(mc, n, 2024, sur_pair=F, sur_pair_sym=B)
    assert r["status"] == 0 and r["n"] == n
    assert (r["stage"] == 2).any() and (r["stage"] == 0).any()
    cur = [0, 0, 0, 0]
    counts = np.zeros(len(states))
    for k in range(n):
        if r["stage"][k] == 1:
            v = 1 if r["fwd"][k] else 0
            cur[r["i"][k]] = v
            cur[r["j"][k]] = v
        counts[states.index(tuple(cur))] += 1
    f = counts / n
    assert np.abs(f - pi).max() < 0.015, (f, pi)


def _lemire(stream, k, n):
    g = int(stream[k]); k += 1
    prod = g * n
    low = prod & 0xFFFFFFFF
    if low < n:
        thr = ((1 << 32) - n) % n
        while low < thr:
            g = int(stream[k]); k += 1
            prod = g * n
            low = prod & 0xFFFFFFFF
    return prod >> 32, k


def test_da_proposal_stream_when_stage1_rejects_everything():
    """With dE* = +1e6 on every trial (mu = 0) stage 1 rejects every move, so
    the pools never change and trial t consumes: a direction canonical (2
    outputs, only when both directions are possible), two Lemire draws, u1 and
    u2 (2 outputs each).  Rebuilt here from the raw mt19937 stream."""
    s, p, _ = S.config_tiny(order=4, grid=(16, 16, 48))
    mc = S.McParams(beta=1.0, mu_eff=0.0)
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    o.nlist()
    n = 400
    r = o.da_run(mc, n, 77, sur_trial=np.full(n, 1e6))
    assert (r["stage"] == 0).all() and np.isnan(r["dE"]).all()
    t = s.type0 + 1
    nh_n = np.flatnonzero(t == S.TYPE_NH_N); nh_p = np.flatnonzero(t == S.TYPE_NH_P)
    zc = s.z[np.concatenate([nh_n, nh_p])].max() + mc.w_z_extra
    w = np.array([k for k in np.flatnonzero(t == S.TYPE_W) if s.z[k] < zc])
    cl = np.array([k for k in np.flatnonzero(t == S.TYPE_CL) if s.z[k] < zc])
    wf, wb = len(nh_n) * len(w), len(nh_p) * len(cl)
    assert wf > 0 and wb > 0
    stream = O.mt_draws(77, 20 * n)
    k = 0
    for m in range(n):
        u = (float(stream[k]) + float(stream[k + 1]) * 4294967296.0) / 18446744073709551616.0
        k += 2
        fwd = u < wf / (wf + wb)
        a, b = (nh_n, w) if fwd else (nh_p, cl)
        ia, k = _lemire(stream, k, len(a))
        ja, k = _lemire(stream, k, len(b))
        k += 4
        assert (r["fwd"][m], r["i"][m], r["j"][m]) == (int(fwd), a[ia], b[ja]), m


# ---------------------------------------------------------------- MLP oracle
def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100),
                        [0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(MLP.bf16(v), ref)


def _theta(seed):
    th = MLP.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    th[:MLP.NTRAIN] += rng.normal(0, 0.05, MLP.NTRAIN)           # non-zero biases too
    o = MLP.NTRAIN
    th[o:o + 48] = rng.uniform(0.05, 0.3, 48)
    th[o + 48], th[o + 49] = 0.7, 2.5
    return th


def test_mlp_forward_matches_torch_fp64():
    th = _theta(1)
    d = np.random.default_rng(2).integers(-20, 20, (64, 48)).astype(np.float64)
    y, _ = MLP.forward(th, d, bf=False)
    W1, b1, W2, b2, w3, b3, sc, om, osd = (torch.from_numpy(np.asarray(a, np.float64)) if np.ndim(a) else float(a)
                                           for a in MLP.unpack(th))
    x = torch.from_numpy(d) * sc
    h1 = torch.nn.functional.gelu(x @ W1 + b1)
    h2 = torch.nn.functional.gelu(h1 @ W2 + b2)
    ref = om + osd * (h2 @ w3 + b3)
    assert np.allclose(y, ref.numpy(), rtol=1e-13, atol=1e-12)


def test_mlp_gradient_finite_differences():
    th = _theta(3)
    rng = np.random.default_rng(4)
    d = rng.integers(-10, 10, (32, 48)).astype(np.float64)
    t = rng.normal(0.5, 2.0, 32)
    loss, g = MLP.loss_and_grad(th, d, t, bf=False)
    idx = np.concatenate([rng.choice(48 * 128, 12, replace=False),
                          48 * 128 + rng.choice(128, 4, replace=False),
                          48 * 128 + 128 + rng.choice(128 * 128, 12, replace=False),
                          [MLP.NTRAIN - 130, MLP.NTRAIN - 5, MLP.NTRAIN - 1]])
    h = 1e-6
    for k in idx:
        tp, tm = th.copy(), th.copy()
        tp[k] += h; tm[k] -= h
        fd = (MLP.loss_and_grad(tp, d, t, bf=False)[0] - MLP.loss_and_grad(tm, d, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_mlp_gradient_matches_torch_autograd():
    th = _theta(5)
    rng = np.random.default_rng(6)
    d = rng.integers(-10, 10, (50, 48)).astype(np.float64)
    t = rng.normal(0.0, 3.0, 50)
    loss, g = MLP.loss_and_grad(th, d, t, bf=False)
    p = torch.tensor(th[:MLP.NTRAIN], dtype=torch.float64, requires_grad=True)
    o = 0
    W1 = p[o:o + 6144].view(48, 128); o += 6144
    b1 = p[o:o + 128]; o += 128
    W2 = p[o:o + 16384].view(128, 128); o += 16384
    b2 = p[o:o + 128]; o += 128
    w3 = p[o:o + 128]; o += 128
    b3 = p[o]
    sc = torch.from_numpy(th[MLP.NTRAIN:MLP.NTRAIN + 48])
    om, osd = th[MLP.NTRAIN + 48], th[MLP.NTRAIN + 49]
    x = torch.from_numpy(d) * sc
    y = torch.nn.functional.gelu(torch.nn.functional.gelu(x @ W1 + b1) @ W2 + b2) @ w3 + b3
    L = torch.mean((y - torch.from_numpy((t - om) / osd)) ** 2)
    L.backward()
    assert abs(L.item() - loss) < 1e-12
    assert np.allclose(p.grad.numpy(), g, rtol=1e-10, atol=1e-13)


def test_adam_matches_torch():
    th = _theta(7)
    rng = np.random.default_rng(8)
    p = torch.tensor(th[:MLP.NTRAIN].copy(), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([p], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(MLP.NTRAIN); v = np.zeros(MLP.NTRAIN)
    cur = th.copy()
    for step in range(1, 4):
        g = rng.normal(0, 1, MLP.NTRAIN)
        p.grad = torch.from_numpy(g.copy())
        opt.step()
        cur, m, v = MLP.adam_step(cur, g, m, v, step, 3e-3)
    assert np.allclose(cur[:MLP.NTRAIN], p.detach().numpy(), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("bf", [False, True])
def test_mlp_bf16_rounding_points(bf):
    """The bf16 forward differs from FP64 by bf16-sized amounts only."""
    th = _theta(9)
    d = np.random.default_rng(10).integers(-20, 20, (128, 48)).astype(np.float64)
    y64, _ = MLP.forward(th, d, bf=False)
    y, _ = MLP.forward(th, d, bf=bf)
    err = np.abs(y - y64).max() / max(1.0, np.abs(y64).max())
    assert (err == 0.0) if not bf else (0.0 < err < 0.05)
