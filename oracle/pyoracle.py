"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU FP64 restatement of the reference's energy path and move rule.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; the product (ph_brush_amd/) never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_pu = C.POINTER(C.c_uint)
_pb = C.POINTER(C.c_byte)


class OrSystem(C.Structure):
    _fields_ = [("natoms", C.c_int), ("x", _pd), ("y", _pd), ("z", _pd), ("q", _pd), ("type", _pi),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double), ("bhead", _pi), ("blist", _pi)]


class OrParams(C.Structure):
    _fields_ = [("ntypes", C.c_int), ("lj_eps", _pd), ("lj_sig", _pd), ("lj_shift", C.c_int), ("rc", C.c_double),
                ("lj_rc", C.c_double), ("epsilon_r", C.c_double), ("coulombtype", C.c_int),
                ("ewald_alpha", C.c_double), ("pme_order", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("nz", C.c_int), ("pme_3dc", C.c_int)]


class OrMcParams(C.Structure):
    _fields_ = [("nh_n_type", C.c_int), ("nh_p_type", C.c_int), ("w_type", C.c_int), ("cl_type", C.c_int),
                ("beta", C.c_double), ("mu_eff", C.c_double), ("w_z_extra", C.c_double)]


class OrRecord(C.Structure):
    _fields_ = [("fwd", _pb), ("i", _pi), ("j", _pi), ("acc", _pb), ("dE", _pd), ("expo", _pd), ("u", _pd)]


class OrDescParams(C.Structure):
    _fields_ = [("species_of_type", _pb), ("ntypes", C.c_int), ("r_desc", C.c_double)]


class OrPhase1Result(C.Structure):
    _fields_ = [("attempted", C.c_longlong), ("accepted", C.c_longlong), ("accepted_forward", C.c_longlong),
                ("accepted_backward", C.c_longlong), ("blocks_done", C.c_int), ("Nh", C.c_int), ("Nprot", C.c_int),
                ("target_count", C.c_int), ("reached_target", C.c_int), ("direct", C.c_int), ("status", C.c_int),
                ("E", C.c_double), ("mu_eff", C.c_double)]


class OrDaRecord(C.Structure):
    _fields_ = [("fwd", _pb), ("i", _pi), ("j", _pi), ("stage", _pb), ("dE_star", _pd), ("dE_star_rev", _pd),
                ("dE", _pd)]


_lib = None


REF_SOURCES = Path("/root/reference")
REF_IO_DUMP = HERE / "_ref" / "ref_io_dump"


def build() -> Path:
    """Compile liboracle.so with the committed Makefile (g++), and, where the
    reference sources are present, oracle/_ref/ref_io_dump from them."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if (REF_SOURCES / "MC" / "src" / "io.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "-B", "ref", f"REF={REF_SOURCES}"], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH))
        lib.or_good_fft_dim.restype = C.c_int
        lib.or_alpha_from_rtol.restype = C.c_double
        lib.or_alpha_from_rtol.argtypes = [C.c_double, C.c_double]
        lib.or_bspline.argtypes = [C.c_int, C.c_double, _pd]
        lib.or_moduli.argtypes = [C.c_int, C.c_int, _pd]
        lib.or_nlist_half.restype = C.c_longlong
        lib.or_nlist_half.argtypes = [C.POINTER(OrSystem), C.c_double, _pi, _pi, C.c_longlong]
        for f in ("or_energy_lj", "or_energy_coul_cut", "or_energy_real_ewald"):
            getattr(lib, f).restype = C.c_double
            getattr(lib, f).argtypes = [C.POINTER(OrSystem), C.POINTER(OrParams), _pi, _pi]
        lib.or_spread.argtypes = [C.POINTER(OrSystem), C.POINTER(OrParams), _pd]
        lib.or_fft_r2c_3d.argtypes = [C.c_int, C.c_int, C.c_int, _pd, _pd]
        lib.or_fft_c2r_3d.argtypes = [C.c_int, C.c_int, C.c_int, _pd, _pd]
        lib.or_influence.argtypes = [C.POINTER(OrParams), C.c_double, C.c_double, C.c_double, _pd]
        lib.or_recip_energy.restype = C.c_double
        lib.or_recip_energy.argtypes = [C.POINTER(OrParams), C.c_double, C.c_double, C.c_double, _pd]
        lib.or_pme_energy.restype = C.c_double
        lib.or_pme_energy.argtypes = [C.POINTER(OrSystem), C.POINTER(OrParams), _pi, _pi, _pd]
        lib.or_total_energy.restype = C.c_double
        lib.or_total_energy.argtypes = [C.POINTER(OrSystem), C.POINTER(OrParams), _pi, _pi, _pd]
        lib.or_gc_run.restype = C.c_int
        lib.or_gc_run.argtypes = [C.POINTER(OrSystem), C.POINTER(OrParams), C.POINTER(OrMcParams), _pi, _pi,
                                  C.c_longlong, C.c_uint, C.POINTER(OrRecord), _pd, _pi, _pd,
                                  C.POINTER(C.c_longlong)]
        lib.or_da_run.restype = C.c_int
        lib.or_da_run.argtypes = [C.POINTER(OrSystem), C.POINTER(OrParams), C.POINTER(OrMcParams), _pi, _pi,
                                  C.c_longlong, C.c_uint, _pd, _pd, _pd, _pd, C.POINTER(OrDaRecord), _pd, _pi,
                                  _pd, C.POINTER(C.c_longlong)]
        lib.or_phase2_run.restype = C.c_int
        lib.or_phase2_run.argtypes = [C.POINTER(OrSystem), C.POINTER(OrParams), C.POINTER(OrMcParams), _pi, _pi,
                                      _pu, _pi, _pi, C.c_longlong, C.c_int, C.c_double, _pi, _pb, _pd, _pd, _pi,
                                      _pd, _pd, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), _pu]
        lib.or_phase1_run.restype = C.c_int
        lib.or_phase1_run.argtypes = [C.POINTER(OrSystem), C.POINTER(OrParams), C.POINTER(OrMcParams), _pi, _pi,
                                      C.c_uint, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_longlong, C.POINTER(OrPhase1Result), _pd, _pi, _pu, _pi, _pi,
                                      C.POINTER(C.c_int)]
        lib.or_desc_thresholds.argtypes = [C.c_double, _pd]
        lib.or_da_descriptors.argtypes = [C.POINTER(OrSystem), C.POINTER(OrDescParams), C.c_int, _pi, _pi, _pi, _pd]
        lib.or_mt_draws.argtypes = [C.c_uint, C.c_longlong, _pu]
        lib.or_uniform_int_draws.argtypes = [C.c_uint, C.c_longlong, _pi, _pi]
        lib.or_canonical_draws.argtypes = [C.c_uint, C.c_longlong, _pd]
        lib.or_mixed_draws.argtypes = [_pu, C.c_longlong, _pi, C.POINTER(C.c_ulonglong), _pu]
        lib.or_set_threads.argtypes = [C.c_int]
        _lib = lib
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_pd)


class Oracle:
    """Holds ctypes views of one System + Params (keeps arrays alive)."""

    def __init__(self, s, p, bonds_csr=None):
        self.lib = load()
        self.s, self.p = s, p
        # private copies: set_charges must never write through to the caller's arrays
        self._arr = [np.array(a, dtype=np.float64, copy=True) for a in (s.x, s.y, s.z, s.q)]
        self._t = np.array(s.type0, dtype=np.int32, copy=True)
        if bonds_csr is not None:
            self._bh = np.ascontiguousarray(bonds_csr[0], np.int32)
            self._bl = np.ascontiguousarray(bonds_csr[1] if len(bonds_csr[1]) else np.zeros(1), np.int32)
            bh, bl = self._bh.ctypes.data_as(_pi), self._bl.ctypes.data_as(_pi)
        else:
            bh, bl = None, None
        self.sys = OrSystem(s.natoms, *(a.ctypes.data_as(_pd) for a in self._arr), self._t.ctypes.data_as(_pi),
                            s.Lx, s.Ly, s.Lz, bh, bl)
        self._eps = np.ascontiguousarray(p.lj_eps, np.float64)
        self._sig = np.ascontiguousarray(p.lj_sig, np.float64)
        nx, ny, nz = p.grid
        self.par = OrParams(p.ntypes, self._eps.ctypes.data_as(_pd), self._sig.ctypes.data_as(_pd), int(p.lj_shift),
                            p.rc, p.lj_rc, p.epsilon_r, p.coulombtype, p.ewald_alpha, p.pme_order, nx, ny, nz,
                            int(p.pme_3dc))
        self.head = None
        self.list = None

    def set_charges(self, q, type0):
        self._arr[3][:] = q
        self._t[:] = type0

    def nlist(self, rc=None):
        rc = self.p.rc if rc is None else rc
        head = np.zeros(self.s.natoms + 1, np.int32)
        nnz = self.lib.or_nlist_half(C.byref(self.sys), rc, head.ctypes.data_as(_pi), None, 0)
        lst = np.zeros(max(nnz, 1), np.int32)
        self.lib.or_nlist_half(C.byref(self.sys), rc, head.ctypes.data_as(_pi), lst.ctypes.data_as(_pi), nnz)
        self.head, self.list = head, lst
        return head, lst[:nnz]

    def _hl(self):
        if self.head is None:
            self.nlist()
        return self.head.ctypes.data_as(_pi), self.list.ctypes.data_as(_pi)

    def energy_lj(self):
        return self.lib.or_energy_lj(C.byref(self.sys), C.byref(self.par), *self._hl())

    def energy_coul_cut(self):
        return self.lib.or_energy_coul_cut(C.byref(self.sys), C.byref(self.par), *self._hl())

    def energy_real(self):
        return self.lib.or_energy_real_ewald(C.byref(self.sys), C.byref(self.par), *self._hl())

    def pme_energy(self):
        c = np.zeros(5)
        e = self.lib.or_pme_energy(C.byref(self.sys), C.byref(self.par), *self._hl(), c.ctypes.data_as(_pd))
        return e, c

    def total_energy(self):
        c = np.zeros(6)
        e = self.lib.or_total_energy(C.byref(self.sys), C.byref(self.par), *self._hl(), c.ctypes.data_as(_pd))
        return e, c

    def spread(self):
        nx, ny, nz = self.p.grid
        g = np.zeros(nx * ny * nz)
        self.lib.or_spread(C.byref(self.sys), C.byref(self.par), g.ctypes.data_as(_pd))
        return g

    def influence(self):
        nx, ny, nz = self.p.grid
        e = np.zeros((nx // 2 + 1) * ny * nz)
        self.lib.or_influence(C.byref(self.par), self.s.Lx, self.s.Ly, self.s.Lz, e.ctypes.data_as(_pd))
        return e

    def gc_run(self, mc, ntrials, seed, record=True):
        """Fixed-mu GC chain with full recompute per trial (the reference algorithm)."""
        N = self.s.natoms
        rec = {k: np.zeros(ntrials, dt) for k, dt in
               (("fwd", np.int8), ("i", np.int32), ("j", np.int32), ("acc", np.int8), ("dE", np.float64),
                ("expo", np.float64), ("u", np.float64))}
        r = OrRecord(*(rec[k].ctypes.data_as(t) for k, t in
                       (("fwd", _pb), ("i", _pi), ("j", _pi), ("acc", _pb), ("dE", _pd), ("expo", _pd), ("u", _pd))))
        m = OrMcParams(mc.nh_n_type, mc.nh_p_type, mc.w_type, mc.cl_type, mc.beta, mc.mu_eff, mc.w_z_extra)
        q, t = np.zeros(N), np.zeros(N, np.int32)
        E = C.c_double()
        nd = C.c_longlong()
        st = self.lib.or_gc_run(C.byref(self.sys), C.byref(self.par), C.byref(m), *self._hl(), ntrials, seed,
                                C.byref(r) if record else None, q.ctypes.data_as(_pd), t.ctypes.data_as(_pi),
                                C.byref(E), C.byref(nd))
        n = nd.value
        return {"status": st, "n": n, "q": q, "type0": t, "E": E.value, **{k: v[:n] for k, v in rec.items()}}

    def phase1_run(self, mc, seed, target_pct, direct_pct_threshold=90, gc_block=2000, epoch_attempt_max=100000,
                   nacc_frac=0.10, mu_eta=0.3, max_trials=0):
        """Phase 1 as main() runs it (build_plan checks, direct flips or the
        mu-control GC loop).  Returns the result fields, the identities, the
        engine state after it and the pools it hands to phase 2."""
        N = self.s.natoms
        m = OrMcParams(mc.nh_n_type, mc.nh_p_type, mc.w_type, mc.cl_type, mc.beta, mc.mu_eff, mc.w_z_extra)
        res = OrPhase1Result()
        q, t = np.zeros(N), np.zeros(N, np.int32)
        rng = np.zeros(625, np.uint32)
        ho, hs = np.zeros(max(N, 1), np.int32), np.zeros(4, np.int32)
        err = C.c_int()
        rc = self.lib.or_phase1_run(C.byref(self.sys), C.byref(self.par), C.byref(m), *self._hl(), seed, target_pct,
                                    direct_pct_threshold, gc_block, epoch_attempt_max, nacc_frac, mu_eta, max_trials,
                                    C.byref(res), q.ctypes.data_as(_pd), t.ctypes.data_as(_pi),
                                    rng.ctypes.data_as(_pu), ho.ctypes.data_as(_pi), hs.ctypes.data_as(_pi),
                                    C.byref(err))
        out = {k: getattr(res, k) for k, _ in OrPhase1Result._fields_}
        out.update(rc=rc, err=err.value, q=q, type0=t, rng_state=rng, handover=ho[:hs.sum()], handover_sizes=hs)
        return out

    def phase2_run(self, mc, rng_state, max_attempts, patience, best_eps=1e-9, pools=None, pool_sizes=None):
        """phase2_swap_jitter restated (MC/src/mc.cu:159-315) from the current
        identities (pools: the phase-1 hand-over, else rebuilt in ascending atom
        order), the engine resumed from rng_state (625 words)."""
        N = self.s.natoms
        st = np.ascontiguousarray(rng_state, dtype=np.uint32)
        assert st.shape == (625,)
        atoms = np.zeros((max_attempts, 4), np.int32)
        acc = np.zeros(max_attempts, np.int8)
        dE = np.zeros(max_attempts)
        m = OrMcParams(mc.nh_n_type, mc.nh_p_type, mc.w_type, mc.cl_type, mc.beta, mc.mu_eff, mc.w_z_extra)
        q, t = np.zeros(N), np.zeros(N, np.int32)
        E, best = C.c_double(), C.c_double()
        after = np.zeros(625, np.uint32)
        att, nacc = C.c_longlong(), C.c_longlong()
        mode, stop = C.c_int(), C.c_int()
        pl = np.ascontiguousarray(pools, np.int32) if pools is not None else None
        ps = np.ascontiguousarray(pool_sizes, np.int32) if pool_sizes is not None else None
        self.lib.or_phase2_run(C.byref(self.sys), C.byref(self.par), C.byref(m), *self._hl(), st.ctypes.data_as(_pu),
                               pl.ctypes.data_as(_pi) if pl is not None else None,
                               ps.ctypes.data_as(_pi) if ps is not None else None, max_attempts, patience, best_eps, atoms.ctypes.data_as(_pi), acc.ctypes.data_as(_pb),
                               dE.ctypes.data_as(_pd), q.ctypes.data_as(_pd), t.ctypes.data_as(_pi), C.byref(E),
                               C.byref(best), C.byref(att), C.byref(nacc), C.byref(mode), C.byref(stop),
                               after.ctypes.data_as(_pu))
        n = att.value
        return {"n": n, "accepted": nacc.value, "mode": mode.value, "stop": stop.value, "E": E.value,
                "best_E": best.value, "q": q, "type0": t, "atoms": atoms[:n], "acc": acc[:n], "dE": dE[:n],
                "rng_state": after}

    def da_descriptors(self, species_of_type, r_desc, i, j):
        """Brute-force shell-count descriptors d = h_i - h_j (int, n x 48) and r_ij."""
        sp = np.ascontiguousarray(species_of_type, dtype=np.int8)
        dp = OrDescParams(sp.ctypes.data_as(_pb), len(sp), float(r_desc))
        i = np.ascontiguousarray(i, dtype=np.int32); j = np.ascontiguousarray(j, dtype=np.int32)
        n = len(i)
        d = np.zeros((n, 48), np.int32)
        rij = np.zeros(n)
        self.lib.or_da_descriptors(C.byref(self.sys), C.byref(dp), n, i.ctypes.data_as(_pi), j.ctypes.data_as(_pi),
                                   d.ctypes.data_as(_pi), rij.ctypes.data_as(_pd))
        return d, rij

    def da_run(self, mc, ntrials, seed, sur_trial=None, sur_rev_trial=None, sur_pair=None, sur_pair_sym=None):
        """Delayed-acceptance chain with the surrogate given per trial (forward and
        reverse values) or as N x N tables (antisymmetric part A, symmetric part B)."""
        N = self.s.natoms
        rec = {k: np.zeros(ntrials, dt) for k, dt in
               (("fwd", np.int8), ("i", np.int32), ("j", np.int32), ("stage", np.int8), ("dE_star", np.float64),
                ("dE_star_rev", np.float64), ("dE", np.float64))}
        r = OrDaRecord(*(rec[k].ctypes.data_as(t) for k, t in
                         (("fwd", _pb), ("i", _pi), ("j", _pi), ("stage", _pb), ("dE_star", _pd),
                          ("dE_star_rev", _pd), ("dE", _pd))))
        m = OrMcParams(mc.nh_n_type, mc.nh_p_type, mc.w_type, mc.cl_type, mc.beta, mc.mu_eff, mc.w_z_extra)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                for a in (sur_trial, sur_rev_trial, sur_pair, sur_pair_sym)]
        q, t = np.zeros(N), np.zeros(N, np.int32)
        E = C.c_double()
        nd = C.c_longlong()
        st = self.lib.or_da_run(C.byref(self.sys), C.byref(self.par), C.byref(m), *self._hl(), ntrials, seed,
                                *(a.ctypes.data_as(_pd) if a is not None else None for a in arrs), C.byref(r),
                                q.ctypes.data_as(_pd), t.ctypes.data_as(_pi), C.byref(E), C.byref(nd))
        n = nd.value
        return {"status": st, "n": n, "q": q, "type0": t, "E": E.value, **{k: v[:n] for k, v in rec.items()}}


def desc_thresholds(r_desc):
    th = np.zeros(12)
    load().or_desc_thresholds(float(r_desc), th.ctypes.data_as(_pd))
    return th


def good_fft_dim(n):
    return load().or_good_fft_dim(n)


def alpha_from_rtol(rc, rtol):
    return load().or_alpha_from_rtol(rc, rtol)


def bspline(order, t):
    w = np.zeros(order)
    load().or_bspline(order, t, w.ctypes.data_as(_pd))
    return w


def moduli(n, order):
    m = np.zeros(n)
    load().or_moduli(n, order, m.ctypes.data_as(_pd))
    return m


def fft_r2c_3d(grid, nx, ny, nz):
    out = np.zeros(2 * (nx // 2 + 1) * ny * nz)
    load().or_fft_r2c_3d(nx, ny, nz, _d(grid), out.ctypes.data_as(_pd))
    return out.view(np.complex128)


def fft_c2r_3d(kgrid, nx, ny, nz):
    k = np.ascontiguousarray(kgrid, np.complex128).view(np.float64)
    out = np.zeros(nx * ny * nz)
    load().or_fft_c2r_3d(nx, ny, nz, k.ctypes.data_as(_pd), out.ctypes.data_as(_pd))
    return out


def mt_draws(seed, n):
    out = np.zeros(n, np.uint32)
    load().or_mt_draws(seed, n, out.ctypes.data_as(_pu))
    return out


def uniform_int_draws(seed, his):
    his = np.ascontiguousarray(his, np.int32)
    out = np.zeros(len(his), np.int32)
    load().or_uniform_int_draws(seed, len(his), his.ctypes.data_as(_pi), out.ctypes.data_as(_pi))
    return out


def canonical_draws(seed, n):
    out = np.zeros(n)
    load().or_canonical_draws(seed, n, out.ctypes.data_as(_pd))
    return out


def mixed_draws(state, his):
    """The proposal stream's mix from an engine state (625 words): his[k] >= 0 draws
    uniform_int_distribution<int>(0, his[k]), his[k] < 0 a canonical double (its
    bits).  Returns the draws (uint64) and the engine state after them."""
    st = np.ascontiguousarray(state, np.uint32)
    assert st.shape == (625,)
    his = np.ascontiguousarray(his, np.int32)
    out = np.zeros(len(his), np.uint64)
    after = np.zeros(625, np.uint32)
    load().or_mixed_draws(st.ctypes.data_as(_pu), len(his), his.ctypes.data_as(_pi),
                          out.ctypes.data_as(C.POINTER(C.c_ulonglong)), after.ctypes.data_as(_pu))
    return out, after


def seed_state(seed):
    """std::mt19937(seed) in its stream form: its 624 words, then the index (624)."""
    st0 = np.zeros(625, np.uint32)
    st0[:624] = _seeded_words(seed)
    st0[624] = 624
    return st0


def _seeded_words(seed):
    # the standard's seeding recurrence x_i = 1812433253 (x_{i-1} ^ (x_{i-1} >> 30)) + i
    w = np.zeros(624, np.uint64)
    w[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        x = int(w[i - 1])
        w[i] = (1812433253 * (x ^ (x >> 30)) + i) & 0xFFFFFFFF
    return w.astype(np.uint32)


def set_threads(n):
    load().or_set_threads(n)
