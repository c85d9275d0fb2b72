"""Pins the CPU oracle (oracle/) against closed forms, numpy, brute force and
known answers, independently of the oracle's own code.

What pins each oracle function (reference file:line of the restated code in
brackets):
  or_mt_draws / or_uniform_int_draws / or_canonical_draws  [MC/src/mc.cu:62-65, :368-369]
      C++ standard known answer (10000th draw of mt19937(5489) = 4123659995) and
      the published Lemire / generate_canonical algorithms applied to the raw stream
  or_bspline                 [MC/src/pme.cu:132-138]  truncated-power form of M_p
  or_moduli                  [MC/src/pme.cu:77-113]   closed forms at orders 4 and 5
  or_good_fft_dim            [MC/src/pme.cu:46-51]    brute-force search
  or_alpha_from_rtol         [MC/src/pme.cu:54-63]    erfc(alpha rc) = rtol by scipy
  or_nlist_half              [MC/src/neighbor.cu:21-125]  numpy brute force, bit-exact
  or_energy_lj / coul / real [MC/src/energy_gpu.cu:14-108, MC/src/pme.cu:252-288]
      two-particle closed forms and a numpy pair sum
  or_spread                  [MC/src/pme.cu:140-172]  charge conservation, dense numpy B-spline sum
  or_fft_r2c_3d / c2r        [MC/src/pme.cu:398-400]  numpy.fft.rfftn / irfftn
  or_influence / or_recip_energy / or_pme_energy [MC/src/pme.cu:175-249, :425-447]
      classical Ewald k-sum, NaCl Madelung constant, alpha invariance, slab limit
  or_gc_run                  [MC/src/mc.cu:379-503]   exact enumeration of a tiny system
"""
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy.special import erfc

from oracle import pyoracle as O
from ph_brush_amd import system as S


# ---------------------------------------------------------------- RNG
def test_mt19937_known_answer():
    # C++ standard [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937 (seed 5489) produces 4123659995.
    assert int(O.mt_draws(5489, 10000)[-1]) == 4123659995


def _lemire(stream, n):
    """Lemire, 'Fast Random Integer Generation in an Interval' (2019), 32-bit."""
    k = 0
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


def test_uniform_int_is_lemire():
    rng = np.random.default_rng(0)
    his = rng.integers(0, 3000, size=4000).astype(np.int32)
    his[:5] = [0, 1, 2, 2**31 - 2, 3]
    got = O.uniform_int_draws(7, his)
    stream = O.mt_draws(7, 20000)
    pos = 0
    for h, v in zip(his, got):
        r, used = _lemire(stream[pos:], int(h) + 1)
        pos += used
        assert r == v


def test_canonical_double_two_draws():
    got = O.canonical_draws(11, 2000)
    stream = O.mt_draws(11, 4000).astype(np.float64)
    # generate_canonical<double,53>: (g1 + g2*2^32) rounded to double, / 2^64
    exp = (stream[0::2] + stream[1::2] * 4294967296.0) / 18446744073709551616.0
    exp = np.where(exp >= 1.0, np.nextafter(1.0, 0.0), exp)
    assert np.array_equal(got, exp)


# ---------------------------------------------------------------- B-splines
def _M(p, u):
    """Cardinal B-spline by the truncated power form (1/(p-1)!) sum (-1)^k C(p,k) (u-k)_+^(p-1),
    evaluated in exact rational arithmetic (the form cancels badly in floating point)."""
    u = Fraction(u)
    v = sum((-1) ** k * math.comb(p, k) * max(u - k, Fraction(0)) ** (p - 1) for k in range(p + 1))
    return float(v / math.factorial(p - 1))


@pytest.mark.parametrize("p", [3, 4, 5, 6, 8])
def test_bspline_weights(p):
    for t in [0.0, 0.1234, 0.5, 0.87, 0.999999]:
        w = O.bspline(p, t)
        exp = np.array([_M(p, t + p - 1 - k) for k in range(p)])
        assert np.allclose(w, exp, atol=1e-13, rtol=0)
        assert abs(w.sum() - 1.0) < 1e-14


def test_moduli_closed_forms():
    for n in [16, 32, 40, 128]:
        th = 2 * np.pi * np.arange(n) / n
        assert np.allclose(O.moduli(n, 4), ((2 + np.cos(th)) / 3) ** 2, rtol=1e-13, atol=0)
        m5 = ((2 * np.cos(1.5 * th) + 22 * np.cos(0.5 * th)) / 24) ** 2
        got = O.moduli(n, 5)
        ok = m5 >= 1e-7
        assert np.allclose(got[ok], m5[ok], rtol=1e-12, atol=0)
        # the near-zero Nyquist modulus is replaced by the mean of its neighbours
        h = n // 2
        assert not ok[h] and abs(got[h] - 0.5 * (m5[h - 1] + m5[h + 1])) < 1e-15


def test_good_fft_dim_bruteforce():
    def smooth(m):
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        return m == 1
    for n in range(1, 300):
        exp = next(m for m in range(max(8, n), 10 * n + 100) if smooth(m))
        assert O.good_fft_dim(n) == exp


def test_alpha_from_rtol():
    for rc, tol in [(1.0, 1e-5), (3.5, 1e-6), (0.9, 1e-4)]:
        a = O.alpha_from_rtol(rc, tol)
        assert abs(erfc(a * rc) - tol) / tol < 1e-9


# ---------------------------------------------------------------- neighbour list
def _nlist_numpy(s, rc):
    bonded = set(map(tuple, np.sort(s.bonds, axis=1).tolist())) if len(s.bonds) else set()
    head, lst = [0], []
    for i in range(s.natoms):
        j = np.arange(i + 1, s.natoms)
        dx = s.x[i] - s.x[j]; dx = dx - s.Lx * np.rint(dx / s.Lx)
        dy = s.y[i] - s.y[j]; dy = dy - s.Ly * np.rint(dy / s.Ly)
        dz = s.z[i] - s.z[j]; dz = dz - s.Lz * np.rint(dz / s.Lz)
        r2 = dx * dx + dy * dy + dz * dz
        hit = [int(k) for k in j[r2 <= rc * rc] if (i, int(k)) not in bonded]
        lst.extend(hit)
        head.append(len(lst))
    return np.array(head, np.int32), np.array(lst, np.int32)


def test_nlist_matches_bruteforce():
    s, p, _ = S.config_tiny()
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    h, l = o.nlist()
    he, le = _nlist_numpy(s, p.rc)
    assert np.array_equal(h, he) and np.array_equal(l, le)
    # rows ascending
    for i in range(s.natoms):
        r = l[h[i]:h[i + 1]]
        assert np.all(np.diff(r) > 0)


# ---------------------------------------------------------------- pair terms
def _two(q1, q2, r, t1=0, t2=0, L=20.0):
    s = S.System(np.array([1.0, 1.0 + r]), np.array([1.0, 1.0]), np.array([1.0, 1.0]), np.array([q1, q2]),
                 np.array([t1, t2], np.int32), np.zeros((0, 2), np.int32), L, L, L, L)
    return s


def test_pair_closed_forms():
    p = S.Params(rc=3.5, lj_rc=2 ** (1 / 6), ewald_alpha=0.9, grid=(16, 16, 16), pme_3dc=False)
    kf = S.KE / p.epsilon_r
    for r in [0.95, 1.0, 2 ** (1 / 6), 1.5, 3.4]:
        o = O.Oracle(_two(1.0, -1.0, r), p)
        # WCA: 4(r^-12 - r^-6) + 1 inside 2^(1/6), zero beyond
        exp_lj = 4 * (r ** -12 - r ** -6) + 1.0 if r <= 2 ** (1 / 6) else 0.0
        assert abs(o.energy_lj() - exp_lj) < 1e-12
        assert abs(o.energy_coul_cut() - kf * (-1.0) / r) < 1e-12
        assert abs(o.energy_real() - kf * (-1.0) * erfc(0.9 * r) / r) < 1e-12
    # E(sigma) = epsilon, E(2^(1/6) sigma) = 0 for the shifted WCA
    assert abs(O.Oracle(_two(0, 0, 1.0), p).energy_lj() - 1.0) < 1e-12
    assert abs(O.Oracle(_two(0, 0, 2 ** (1 / 6)), p).energy_lj()) < 1e-12
    # reference behaviour (no separate LJ cutoff): LJ to rc with shift at rc
    p2 = S.Params(rc=2.5, lj_rc=0.0, grid=(16, 16, 16))
    exp = 4 * (1.2 ** -12 - 1.2 ** -6) - 4 * (2.5 ** -12 - 2.5 ** -6)
    assert abs(O.Oracle(_two(0, 0, 1.2), p2).energy_lj() - exp) < 1e-12


def test_pair_sums_match_numpy():
    s, p, _ = S.config_tiny()
    p.lj_sig = np.ones(64); p.lj_sig[3 * 8 + 7] = p.lj_sig[7 * 8 + 3] = 1.1   # make one pair type differ
    o = O.Oracle(s, p, S.bond_adjacency(s.natoms, s.bonds))
    h, l = o.nlist()
    i = np.repeat(np.arange(s.natoms), np.diff(h))
    j = l
    d = [a[i] - a[j] for a in (s.x, s.y, s.z)]
    d = [dd - L * np.rint(dd / L) for dd, L in zip(d, (s.Lx, s.Ly, s.Lz))]
    r = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    kf = S.KE / p.epsilon_r
    qq = s.q[i] * s.q[j]
    assert abs(o.energy_real() - np.sum(kf * qq * erfc(p.ewald_alpha * r) / r)) < 1e-9
    sig = p.lj_sig.reshape(8, 8)[s.type0[i], s.type0[j]]
    m = r <= p.lj_rc
    e = 4 * ((sig / r) ** 12 - (sig / r) ** 6) - 4 * ((sig / p.lj_rc) ** 12 - (sig / p.lj_rc) ** 6)
    assert abs(o.energy_lj() - np.sum(e[m])) < 1e-8 * max(1.0, abs(np.sum(e[m])))


# ---------------------------------------------------------------- SPME pieces
@pytest.mark.parametrize("dims", [(16, 8, 32), (12, 10, 15), (32, 32, 128), (9, 6, 25)])
def test_fft_matches_numpy(dims):
    nx, ny, nz = dims
    g = np.random.default_rng(1).standard_normal((nz, ny, nx))
    k = O.fft_r2c_3d(g.ravel(), nx, ny, nz).reshape(nz, ny, nx // 2 + 1)
    ref = np.fft.rfftn(g)
    assert np.abs(k - ref).max() < 1e-11 * np.abs(ref).max()
    back = O.fft_c2r_3d(ref.ravel(), nx, ny, nz).reshape(nz, ny, nx)
    exp = np.fft.irfftn(ref, s=(nz, ny, nx), axes=(0, 1, 2)) * (nx * ny * nz)
    assert np.abs(back - exp).max() < 1e-11 * np.abs(exp).max()


def test_spread_matches_dense_bspline():
    s, p, _ = S.config_tiny(order=5, grid=(12, 10, 30))
    o = O.Oracle(s, p)
    g = o.spread().reshape(30, 10, 12)
    assert abs(g.sum() - s.q.sum()) < 1e-10
    # dense evaluation: weight of grid point m for an atom is M_p(u - m + s), s = p//2
    P = p.pme_order
    exp = np.zeros_like(g)
    for a in range(s.natoms):
        if s.q[a] == 0:
            continue
        ws = []
        for v, L, n in ((s.x[a], s.Lx, 12), (s.y[a], s.Ly, 10), (s.z[a], s.Lz, 30)):
            u = v / L * n
            w = np.zeros(n)
            for m in range(int(np.floor(u)) - P, int(np.floor(u)) + P + 1):
                w[m % n] += _M(P, u - m + P // 2)
            ws.append(w)
        exp += s.q[a] * np.einsum("z,y,x->zyx", ws[2], ws[1], ws[0])
    assert np.abs(g - exp).max() < 1e-12


def _ewald_k_sum(s, alpha, lB, mmax):
    ch = s.q != 0
    x, y, z, q = s.x[ch], s.y[ch], s.z[ch], s.q[ch]
    L = np.array([s.Lx, s.Ly, s.Lz]); V = L.prod()
    E = 0.0
    r = [np.arange(-m, m + 1) for m in mmax]
    for kz in r[2]:
        for ky in r[1]:
            kx = r[0]
            kv = 2 * np.pi * np.stack([kx / L[0], np.full(kx.shape, ky) / L[1], np.full(kx.shape, kz) / L[2]], 1)
            k2 = (kv ** 2).sum(1)
            ok = k2 > 0
            ph = np.outer(kv[:, 0], x) + np.outer(kv[:, 1], y) + np.outer(kv[:, 2], z)
            Sk = (q * np.exp(1j * ph)).sum(1)
            E += np.sum((4 * np.pi / k2[ok]) * np.exp(-k2[ok] / (4 * alpha ** 2)) * np.abs(Sk[ok]) ** 2)
    return E * lB / (2 * V)


def _random_neutral(n, L, seed):
    rng = np.random.default_rng(seed)
    pos = []
    while len(pos) < n:
        c = rng.uniform(0, L, 3)
        if all(np.linalg.norm((c - p_) - L * np.rint((c - p_) / L)) > 0.9 for p_ in pos):
            pos.append(c)
    pos = np.array(pos)
    q = np.array([1.0, -1.0] * (n // 2))
    return S.System(pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), q, np.zeros(n, np.int32),
                    np.zeros((0, 2), np.int32), L, L, L, L)


@pytest.mark.parametrize("order", [4, 5, 6])
def test_spme_recip_vs_classical_ewald(order):
    s = _random_neutral(24, 6.0, 3)
    p = S.Params(rc=2.9, ewald_alpha=1.0, pme_order=order, grid=(40, 40, 40), pme_3dc=False,
                 lj_eps=np.zeros(64))
    o = O.Oracle(s, p)
    _, c = o.pme_energy()
    exp = _ewald_k_sum(s, 1.0, 3.0, (8, 8, 8))
    tol = {4: 2e-4, 5: 3e-5, 6: 5e-6}[order]
    assert abs(c[1] - exp) < tol * abs(exp)


def test_ewald_alpha_invariance():
    s = _random_neutral(30, 8.0, 5)
    tot = []
    for a in (0.95, 1.15):
        p = S.Params(rc=3.95, ewald_alpha=a, pme_order=6, grid=(48, 48, 48), pme_3dc=False, lj_eps=np.zeros(64))
        e, c = O.Oracle(s, p).pme_energy()
        tot.append(e)
    assert abs(tot[0] - tot[1]) < 2e-5 * abs(tot[0])


def test_nacl_madelung():
    n, d = 4, 1.0
    idx = np.array([(i, j, k) for i in range(n) for j in range(n) for k in range(n)], float)
    q = np.where(idx.sum(1) % 2 == 0, 1.0, -1.0)
    L = n * d
    s = S.System(idx[:, 0] * d, idx[:, 1] * d, idx[:, 2] * d, q, np.zeros(len(q), np.int32),
                 np.zeros((0, 2), np.int32), L, L, L, L)
    p = S.Params(rc=1.95, ewald_alpha=2.6, pme_order=6, grid=(48, 48, 48), pme_3dc=False, lj_eps=np.zeros(64))
    e, _ = O.Oracle(s, p).pme_energy()
    M = 1.747564594633182
    assert abs(e / (len(q) / 2) - (-M * 3.0 / d)) < 1e-6 * M * 3.0


def test_slab_correction_large_vacuum_limit():
    base, p, _ = S.config_tiny(order=5, grid=(16, 16, 48))
    vals = []
    for zfac, nz in ((3.0, 48), (6.0, 96)):
        s = base.copy()
        s.Lz = s.Lz_data * zfac
        p2 = S.Params(rc=p.rc, pme_order=6, grid=(16, 16, nz), pme_3dc=True)
        vals.append(O.Oracle(s, p2, S.bond_adjacency(s.natoms, s.bonds)).total_energy()[0])
    assert abs(vals[0] - vals[1]) < 1e-3 * abs(vals[0])


# ---------------------------------------------------------------- MC move rule
def test_gc_chain_samples_exact_distribution():
    """4 atoms: two sites, two ion slots; states conserve N_P - N_Cl, so from
    all-neutral the chain visits 6 states with pi ~ exp(-beta(E - mu N_P))."""
    x = np.array([1.0, 2.6, 1.2, 3.0]); y = np.array([1.0, 1.3, 2.5, 2.4]); z = np.array([1.0, 1.5, 1.0, 2.2])
    t0 = np.array([S.TYPE_NH_N, S.TYPE_NH_N, S.TYPE_W, S.TYPE_W], np.int32) - 1
    s = S.System(x, y, z, np.zeros(4), t0, np.zeros((0, 2), np.int32), 6.0, 6.0, 6.0, 6.0)
    p = S.Params(rc=2.9, ewald_alpha=0.9, pme_order=4, grid=(16, 16, 16), pme_3dc=False)
    mc = S.McParams(beta=1.0, mu_eff=0.3)
    states = []
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                for d in (0, 1):
                    if a + b == c + d:
                        states.append((a, b, c, d))
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
    n = 60000
    r = o.gc_run(mc, n, 12345)
    assert r["status"] == 0 and r["n"] == n
    cur = [0, 0, 0, 0]
    counts = np.zeros(len(states))
    for k in range(n):
        if r["acc"][k]:
            v = 1 if r["fwd"][k] else 0
            cur[r["i"][k]] = v
            cur[r["j"][k]] = v
        counts[states.index(tuple(cur))] += 1
    f = counts / n
    assert np.abs(f - pi).max() < 0.015, (f, pi)


def test_mixed_draws_from_a_state():
    """The proposal stream's mix (ints with Lemire rejections, canonical doubles)
    from an engine state, against the raw stream and the published algorithms;
    the state form round-trips (a state taken mid-stream resumes the stream)."""
    rng = np.random.default_rng(3)
    n = 3000
    his = rng.integers(0, 5000, size=n).astype(np.int32)
    his[rng.random(n) < 0.3] = -1                      # canonical doubles
    his[rng.random(n) < 0.2] = 1431655765              # 2^32/3 + 1 values: rejection ~1/3 of draws
    his[:4] = [2 ** 31 - 1, -1, 0, 1431655765]
    st0 = O.seed_state(19)
    got, after = O.mixed_draws(st0, his)
    stream = O.mt_draws(19, 3 * n + 100)
    pos = 0
    nrej = 0
    for h, v in zip(his, got):
        if h >= 0:
            r, used = _lemire(stream[pos:], int(h) + 1)
            nrej += used - 1
            pos += used
            assert r == v
        else:
            g1, g2 = float(stream[pos]), float(stream[pos + 1])
            pos += 2
            e = (g1 + g2 * 4294967296.0) / 18446744073709551616.0
            e = min(e, np.nextafter(1.0, 0.0))
            assert np.float64(e).view(np.uint64) == v
    assert nrej > 100
    # resuming from the state after the draws continues the raw stream
    more, _ = O.mixed_draws(after, np.full(5, 2 ** 31 - 1, np.int32))
    exp = [_lemire(stream[pos + k:], 2 ** 31)[0] for k in range(5)]
    assert list(more) == exp


def test_seed_state_is_the_standard_seeding():
    """std::mt19937(seed)'s stream form: the standard's seeding recurrence, index 624;
    the C ABI's host helper gives the same words."""
    from ph_brush_amd import system as S
    st = O.seed_state(5489)
    assert st[624] == 624 and st[0] == 5489 and st[1] == 1301868182
    assert np.array_equal(S.rng_seed_state(5489), st)
    assert np.array_equal(S.rng_seed_state(2025), O.seed_state(2025))
