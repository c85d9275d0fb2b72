"""Training data, training and evaluation around the delayed-acceptance chain.

No reference counterpart (the reference has no learned part).  Everything
here drives the C ABI: descriptors, exact Delta-E and training run on the
GPU; the host only draws candidate moves and sequences the calls.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import system as S


def move_pools(type0: np.ndarray, z: np.ndarray, mc: S.McParams):
    """Sites and ions eligible for a move, as build_plan selects them
    (MC/src/mc.cu:67-100): ions below max site z + w_z_extra."""
    t = type0 + 1
    nh_n = np.flatnonzero(t == mc.nh_n_type)
    nh_p = np.flatnonzero(t == mc.nh_p_type)
    sites = np.concatenate([nh_n, nh_p])
    zc = z[sites].max() + mc.w_z_extra if len(sites) else -np.inf
    w = np.flatnonzero((t == mc.w_type) & (z < zc))
    cl = np.flatnonzero((t == mc.cl_type) & (z < zc))
    return nh_n, nh_p, w, cl


def sample_moves(rng: np.random.Generator, type0, z, mc: S.McParams, n: int):
    """n candidate moves drawn like the chain's proposals: direction with
    probability proportional to its number of (site, ion) pairs, then a
    uniform site and ion of that direction."""
    nh_n, nh_p, w, cl = move_pools(type0, z, mc)
    wf, wb = len(nh_n) * len(w), len(nh_p) * len(cl)
    if wf + wb == 0:
        raise ValueError("no valid move")
    fwd = (rng.random(n) < wf / (wf + wb)).astype(np.int32)
    i = np.where(fwd == 1, nh_n[rng.integers(0, max(len(nh_n), 1), n)] if len(nh_n) else -1,
                 nh_p[rng.integers(0, max(len(nh_p), 1), n)] if len(nh_p) else -1)
    j = np.where(fwd == 1, w[rng.integers(0, max(len(w), 1), n)] if len(w) else -1,
                 cl[rng.integers(0, max(len(cl), 1), n)] if len(cl) else -1)
    return i.astype(np.int32), j.astype(np.int32), fwd


@dataclass
class CollectSpec:
    ph_values: tuple = (6.0, 6.5, 7.0, 7.5, 8.0, 8.5, 9.0)
    pka: float = 7.5
    samples_per_ph: int = 65536
    per_state: int = 4096          # candidate moves evaluated per sampled state
    equilibrate: int = 20000       # plain trials before the first sample at each pH
    between: int = 2000            # plain trials between sampled states
    seed: int = 1
    r_desc: float = 6.0
    real_space: bool = False       # physics part with the neighbours' real-space potential (da_attach)
    local_pairs: bool = False      # local-pair physics part (da_attach); the far field frozen after equilibration


def collect(s: S.System, p: S.Params, sur: S.Surrogate, spec: CollectSpec, log=None) -> dict:
    """Fills the surrogate's dataset with (descriptor, s dE + elF/r_ij) pairs
    from exact chains at each pH of the spec (exact dE from the device)."""
    rng = np.random.default_rng(spec.seed)
    t0 = time.perf_counter()
    n0 = sur.size()
    # one chain visits the pH values in turn (its mu set per pH), equilibrating at each
    rep = S.Replica(s, p, S.McParams.at_ph(spec.ph_values[0], spec.pka), seed=spec.seed)
    rep.mc.da_attach(r_desc=spec.r_desc, real_space=spec.real_space, local_pairs=spec.local_pairs)
    for k, ph in enumerate(spec.ph_values):
        mc = S.McParams.at_ph(ph, spec.pka)
        rep.mc.set_mu(mc.mu_eff)
        rep.mc.run(spec.equilibrate)
        if spec.local_pairs:
            rep.mc.da_freeze()   # the states sampled below see this field, as a chain sees its start's
        got = 0
        while got < spec.samples_per_ph:
            rep.mc.run(spec.between)
            _, t = rep.atoms.identity()
            n = min(spec.per_state, spec.samples_per_ph - got)
            i, j, f = sample_moves(rng, t, s.z, mc, n)
            rep.mc.da_collect(sur, i, j, f)
            got += n
        if log:
            log(f"[surrogate] pH {ph}: {got} samples ({time.perf_counter() - t0:.1f}s)")
    del rep
    return {"samples": sur.size() - n0, "seconds": time.perf_counter() - t0}


def train_holdout(sur: S.Surrogate, epochs: int = 20, batch: int = 4096, lr: float = 2e-3, holdout: float = 0.1,
                  seed: int = 1) -> dict:
    """Trains on a shuffled split: the dataset is held in collection order, so
    the hold-out is every k-th sample moved to the end on the host first."""
    n = sur.size()
    d, t = sur.data()
    perm = np.random.default_rng(seed).permutation(n)
    d, t = d[perm], t[perm]
    sur.clear()
    sur.add(d, t)
    n_train = int(n * (1.0 - holdout)) // batch * batch
    st = sur.train(epochs, batch=batch, lr=lr, seed=seed, n_train=n_train, normalize=True)
    rmse_tr, _ = sur.eval(0, n_train)
    rmse_ho, mae_ho = sur.eval(n_train, n - n_train)
    st.update({"n_train": n_train, "n_holdout": n - n_train, "rmse_train": rmse_tr, "rmse_holdout": rmse_ho,
               "mae_holdout": mae_ho, "target_std": float(np.std(t[n_train:]))})
    return st
