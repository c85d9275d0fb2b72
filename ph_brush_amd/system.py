"""Host-side mirror of the reference's setup path and device objects.

Names follow the reference's host API (MC/include/*.hpp): DeviceAtoms,
DeviceCSR, DeviceParams, build_half_neighbors_gpu, compute_lj_sr_energy_gpu,
compute_coul_sr_energy_gpu, PMEPlan (pmeCreate / pmeEnergy / pmeDestroy).
Each wraps one handle of the C ABI and frees it when garbage-collected.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

KE = 138.935458  # kJ mol^-1 nm e^-2 (MC/src/pme.cu:305)

# type ids of the synthetic brushes (1-based, like the reference's Params;
# NH_N/NH_P/W/Cl are the reference defaults, MC/include/types.hpp:51-54)
TYPE_PERM, TYPE_CATION, TYPE_NH_N, TYPE_NH_P, TYPE_W, TYPE_CL = 1, 2, 4, 5, 7, 8
NTYPES = 8

# descriptor species of the surrogate: 0 charged bead, 1 neutral titratable
# bead, 2 cation, 3 anion; -1 not counted (neutral placeholders)
SPECIES_CHARGED, SPECIES_NEUTRAL_SITE, SPECIES_CATION, SPECIES_ANION = 0, 1, 2, 3


def default_species(ntypes: int = NTYPES) -> np.ndarray:
    """species_of_type (0-based type index) for the synthetic brushes."""
    sp = np.full(ntypes, -1, np.int8)
    sp[TYPE_PERM - 1] = SPECIES_CHARGED
    sp[TYPE_NH_P - 1] = SPECIES_CHARGED
    sp[TYPE_NH_N - 1] = SPECIES_NEUTRAL_SITE
    sp[TYPE_CATION - 1] = SPECIES_CATION
    sp[TYPE_CL - 1] = SPECIES_ANION
    return sp


@dataclass
class System:
    """Host system: wrapped positions, charges, 0-based types, bonds, box."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    q: np.ndarray
    type0: np.ndarray
    bonds: np.ndarray          # (nb, 2) int32, 0-based atom indices
    Lx: float
    Ly: float
    Lz: float                  # simulation box (extended under the slab correction)
    Lz_data: float = 0.0       # data-file box height
    mol: np.ndarray | None = None

    @property
    def natoms(self) -> int:
        return int(self.x.shape[0])

    def copy(self) -> "System":
        return System(self.x.copy(), self.y.copy(), self.z.copy(), self.q.copy(), self.type0.copy(),
                      self.bonds.copy(), self.Lx, self.Ly, self.Lz, self.Lz_data,
                      None if self.mol is None else self.mol.copy())


def bond_adjacency(natoms: int, bonds: np.ndarray):
    """Symmetric 1-2 bond CSR, rows sorted and unique (MC/src/io.cpp:148-169)."""
    rows = [[] for _ in range(natoms)]
    for i, j in np.asarray(bonds, dtype=np.int64).reshape(-1, 2):
        rows[i].append(int(j))
        rows[j].append(int(i))
    head = np.zeros(natoms + 1, dtype=np.int32)
    lst = []
    for i in range(natoms):
        r = sorted(set(rows[i]))
        head[i + 1] = head[i] + len(r)
        lst.extend(r)
    return head, np.asarray(lst, dtype=np.int32)


def wrap_positions(v: np.ndarray, L: float) -> np.ndarray:
    """MC/src/io.cpp:303-316: w = r - floor(r/L)*L, folded into [0, L)."""
    w = v - np.floor(v / L) * L
    w = np.where(w < 0, w + L, w)
    return np.where(w >= L, w - L, w)


def gen_brush(chains_x, chains_y, chain_len, Lx, Ly, Lz, *, spacing=2.0, n_perm=0, p_prot=0.5, n_salt=0,
              bond=0.97, min_dist=0.8, seed=1, graft_z=0.0, zfac=None, types=None) -> System:
    """Synthetic brush from the C generator (phb_gen_brush).  types: 1-based ids
    (perm, cation, nh_n, nh_p, w, cl); default the synthetic brushes' TYPE_*."""
    tp = types or (TYPE_PERM, TYPE_CATION, TYPE_NH_N, TYPE_NH_P, TYPE_W, TYPE_CL)
    spec = L.BrushSpec(chains_x, chains_y, spacing, graft_z, chain_len, n_perm, p_prot, n_salt, Lx, Ly, Lz, bond,
                       min_dist, seed, *tp)
    n = C.c_int(0)
    nb = C.c_int(0)
    null_d = C.POINTER(C.c_double)()
    null_i = C.POINTER(C.c_int)()
    L.call("phb_gen_brush", C.byref(spec), 0, null_d, null_d, null_d, null_d, null_i, null_i, null_i, null_i, 0,
           C.byref(n), C.byref(nb))
    N, NB = n.value, nb.value
    x, y, z, q = (np.zeros(N) for _ in range(4))
    t1, mol = np.zeros(N, np.int32), np.zeros(N, np.int32)
    bi, bj = np.zeros(max(NB, 1), np.int32), np.zeros(max(NB, 1), np.int32)
    L.call("phb_gen_brush", C.byref(spec), N, L.dptr(x), L.dptr(y), L.dptr(z), L.dptr(q), L.iptr(t1), L.iptr(mol),
           L.iptr(bi), L.iptr(bj), max(NB, 1), C.byref(n), C.byref(nb))
    x, y, z = wrap_positions(x, Lx), wrap_positions(y, Ly), wrap_positions(z, Lz)
    bonds = np.stack([bi[:NB], bj[:NB]], axis=1).astype(np.int32)
    Lz_sim = Lz * (max(2.0, zfac) if zfac else 1.0)   # MC/src/main.cu:110-117
    return System(x, y, z, q, (t1 - 1).astype(np.int32), bonds, Lx, Ly, Lz_sim, Lz, mol)


@dataclass
class Params:
    """Physics parameters (params.in keys of MC/src/io.cpp:184-256 plus extensions)."""
    rc: float = 3.5
    lj_rc: float = 2.0 ** (1.0 / 6.0)
    lj_shift: bool = True
    epsilon_r: float = KE / 3.0        # Bjerrum length 3 sigma at beta = 1
    coulombtype: int = L.PHB_COUL_PME
    ewald_alpha: float = 0.9
    ewald_rtol: float = 1e-5
    pme_spacing: float = 0.12
    pme_order: int = 5
    grid: tuple = (32, 32, 128)
    pme_3dc: bool = True
    ntypes: int = NTYPES
    lj_eps: np.ndarray = field(default_factory=lambda: np.ones(NTYPES * NTYPES))
    lj_sig: np.ndarray = field(default_factory=lambda: np.ones(NTYPES * NTYPES))

    def desc(self) -> L.ParamsDesc:
        self._eps = np.ascontiguousarray(self.lj_eps, dtype=np.float64)
        self._sig = np.ascontiguousarray(self.lj_sig, dtype=np.float64)
        nx, ny, nz = self.grid if self.grid else (0, 0, 0)
        return L.ParamsDesc(self.ntypes, L.dptr(self._eps), L.dptr(self._sig), int(self.lj_shift), self.rc,
                            self.lj_rc, self.epsilon_r, self.coulombtype, self.ewald_alpha, self.ewald_rtol,
                            self.pme_spacing, self.pme_order, nx, ny, nz, int(self.pme_3dc))


@dataclass
class McParams:
    """Move parameters (MC/include/types.hpp:51-68 plus fixed-mu extension)."""
    nh_n_type: int = TYPE_NH_N
    nh_p_type: int = TYPE_NH_P
    w_type: int = TYPE_W
    cl_type: int = TYPE_CL
    beta: float = 1.0
    mu_eff: float = 0.0
    w_z_extra: float = 1.0e6

    @staticmethod
    def at_ph(ph: float, pka: float = 7.5, beta: float = 1.0, **kw) -> "McParams":
        """Fixed mu_eff = kT ln10 (pKa - pH)."""
        return McParams(beta=beta, mu_eff=math.log(10.0) * (pka - ph) / beta, **kw)


# ------------------------------------------------------------------ workloads
def config_small_fp64(seed: int = 1):
    """64 chains x 20 beads, 16x16x40 box (x3 slab), every bead titratable at
    pH = pKa = 7.5 with p = 0.5, 40 salt pairs; SPME 32x32x128, order 5."""
    s = gen_brush(8, 8, 20, 16.0, 16.0, 40.0, n_perm=0, p_prot=0.5, n_salt=40, seed=seed, zfac=3.0)
    return s, Params(grid=(32, 32, 128)), McParams.at_ph(7.5)


def config_30k_fp64(seed: int = 1):
    """400 chains x 50 beads on a 20x20 lattice, 40x40x150 box (x3 slab); beads
    1-25 permanently +1, 26-50 titratable (pKa 7.5, pH 7.0); 2,000 salt pairs;
    SPME 64x64x256, order 5."""
    s = gen_brush(20, 20, 50, 40.0, 40.0, 150.0, n_perm=25, p_prot=0.5, n_salt=2000, seed=seed, zfac=3.0)
    return s, Params(grid=(64, 64, 256)), McParams.at_ph(7.0)


def config_titration_grid(seed: int = 7, ph_values=(5.0, 5.5, 6.0, 6.5, 7.0, 7.5, 8.0, 8.5),
                          salts=(200, 1000, 3000, 6000), n_min: int = 5000, n_max: int = 35000):
    """Replicas of a titration grid: every (pH, salt pairs) combination with its
    own brush -- chains_x, chains_y in [10, 20] (100-400 chains) on a 2-sigma
    lattice and chain length in [20, 60], drawn from numpy's default_rng(seed)
    and redrawn until the replica has n_min..n_max atoms.  Every bead titratable
    (pKa 7.5, initially protonated with p = 0.5), one counter-particle per bead;
    box 2*chains_x x 2*chains_y x 2.5*length (x3 slab extension); SPME grid 32 or
    64 laterally and 128 or 256 along z by box size, order 5.  Returns a list of
    (System, Params, McParams, label)."""
    rng = np.random.default_rng(seed)
    out = []
    for salt in salts:
        for ph in ph_values:
            while True:
                cx, cy, ln = int(rng.integers(10, 21)), int(rng.integers(10, 21)), int(rng.integers(20, 61))
                n = 2 * cx * cy * ln + 2 * salt
                if n_min <= n <= n_max:
                    break
            Lx, Ly, Lz = 2.0 * cx, 2.0 * cy, 2.5 * ln
            s = gen_brush(cx, cy, ln, Lx, Ly, Lz, n_perm=0, p_prot=0.5, n_salt=salt, seed=int(rng.integers(1, 2**31)),
                          zfac=3.0)
            grid = (32 if Lx <= 24 else 64, 32 if Ly <= 24 else 64, 128 if 3.0 * Lz <= 200 else 256)
            out.append((s, Params(grid=grid), McParams.at_ph(ph),
                        f"pH {ph} salt {salt}: {cx}x{cy} chains x {ln}, {s.natoms} atoms, grid {grid}"))
    return out


def config_tiny(seed: int = 3, order: int = 4, grid=(16, 16, 48), chains=(3, 3), chain_len=6, n_perm=0,
                n_salt=6, box=(6.0, 6.0, 16.0)):
    """A few hundred atoms for quick parity checks."""
    s = gen_brush(chains[0], chains[1], chain_len, box[0], box[1], box[2], n_perm=n_perm, p_prot=0.5,
                  n_salt=n_salt, seed=seed, zfac=3.0)
    return s, Params(grid=grid, pme_order=order), McParams.at_ph(7.5)


# ------------------------------------------------------------ device objects
class _Handle:
    _destroy = ""

    def __init__(self):
        self.h = C.c_void_p()

    def __del__(self):
        try:
            if self.h:
                getattr(L.load(), self._destroy)(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class DeviceAtoms(_Handle):
    """Replaces DeviceAtoms + upload_atoms_to_device (MC/src/main.cu:18-41)."""
    _destroy = "phb_atoms_destroy"

    def __init__(self, s: System):
        super().__init__()
        self.n = s.natoms
        self._keep = [np.ascontiguousarray(a) for a in (s.x, s.y, s.z, s.q)]
        t = np.ascontiguousarray(s.type0, dtype=np.int32)
        L.call("phb_atoms_create", self.n, *(L.dptr(a) for a in self._keep), L.iptr(t), C.byref(self.h))

    def set_identity(self, idx, type0, q):
        idx = np.ascontiguousarray(idx, np.int32)
        type0 = np.ascontiguousarray(type0, np.int32)
        q = np.ascontiguousarray(q, np.float64)
        L.call("phb_atoms_set_identity", self.h, len(idx), L.iptr(idx), L.iptr(type0), L.dptr(q))

    def identity(self):
        q, t = np.zeros(self.n), np.zeros(self.n, np.int32)
        L.call("phb_atoms_get_identity", self.h, L.dptr(q), L.iptr(t))
        return q, t

    def set_positions(self, x, y, z):
        """New wrapped positions (phb_atoms_set_positions)."""
        v = [np.ascontiguousarray(a, np.float64) for a in (x, y, z)]
        assert all(len(a) == self.n for a in v)
        L.call("phb_atoms_set_positions", self.h, *(L.dptr(a) for a in v))


class DeviceCSR(_Handle):
    """Replaces DeviceCSR / copy_csr_to_device (MC/src/neighbor.cu:65-72)."""
    _destroy = "phb_csr_destroy"

    @classmethod
    def from_host(cls, head, lst):
        c = cls()
        head = np.ascontiguousarray(head, np.int32)
        lst = np.ascontiguousarray(lst if len(lst) else np.zeros(1, np.int32), np.int32)
        L.call("phb_csr_create", len(head) - 1, L.iptr(head), L.iptr(lst), C.byref(c.h))
        return c

    def shape(self):
        r, nnz = C.c_int(0), C.c_longlong(0)
        L.call("phb_csr_shape", self.h, C.byref(r), C.byref(nnz))
        return r.value, nnz.value

    def download(self):
        rows, nnz = self.shape()
        head = np.zeros(rows + 1, np.int32)
        lst = np.zeros(max(nnz, 1), np.int32)
        L.call("phb_csr_download", self.h, L.iptr(head), L.iptr(lst))
        return head, lst[:nnz]


def box_of(s: System) -> L.Box:
    return L.Box(s.Lx, s.Ly, s.Lz)


def build_half_neighbors_gpu(atoms: DeviceAtoms, box: L.Box, bonds12: DeviceCSR | None, rc: float) -> DeviceCSR:
    """MC/include/neighbor.hpp:10-13."""
    c = DeviceCSR()
    L.call("phb_nlist_build_half", atoms.h, C.byref(box), bonds12.h if bonds12 else None, rc, C.byref(c.h))
    return c


class VerletList(_Handle):
    """Skin list (phb_verlet_*): the half list at rc + skin, rebuilt only when an
    atom has moved more than skin / 2 since the last build; half() gives the rc
    list, equal to build_half_neighbors_gpu at the current positions."""
    _destroy = "phb_verlet_destroy"

    def __init__(self, atoms: DeviceAtoms, box: L.Box, bonds12: DeviceCSR | None, rc: float, skin: float = 0.3):
        super().__init__()
        self._keep = (bonds12,)
        L.call("phb_verlet_create", atoms.h, C.byref(box), bonds12.h if bonds12 else None, rc, skin,
               C.byref(self.h))

    def update(self, atoms: DeviceAtoms):
        rebuilt, disp = C.c_int(0), C.c_double(0)
        L.call("phb_verlet_update", self.h, atoms.h, C.byref(rebuilt), C.byref(disp))
        return bool(rebuilt.value), disp.value

    def half(self, atoms: DeviceAtoms) -> DeviceCSR:
        c = DeviceCSR()
        L.call("phb_verlet_half", self.h, atoms.h, C.byref(c.h))
        return c

    def stats(self):
        b, n, d = C.c_longlong(0), C.c_longlong(0), C.c_double(0)
        L.call("phb_verlet_stats", self.h, C.byref(b), C.byref(n), C.byref(d))
        return {"builds": b.value, "outer_nnz": n.value, "last_disp": d.value}


class DeviceParams(_Handle):
    """Replaces DeviceParams + upload_params_to_device (MC/src/main.cu:43-66)."""
    _destroy = "phb_params_destroy"

    def __init__(self, p: Params):
        super().__init__()
        self.p = p
        d = p.desc()
        L.call("phb_params_create", C.byref(d), C.byref(self.h))


def compute_lj_sr_energy_gpu(atoms, params, neigh, box) -> float:
    e = C.c_double(0)
    L.call("phb_energy_lj", atoms.h, params.h, neigh.h, C.byref(box), C.byref(e))
    return e.value


def compute_coul_sr_energy_gpu(atoms, params, neigh, box) -> float:
    e = C.c_double(0)
    L.call("phb_energy_coul_cut", atoms.h, params.h, neigh.h, C.byref(box), C.byref(e))
    return e.value


class PMEPlan(_Handle):
    """pmeCreate / pmeDestroy / pmeEnergy (MC/include/pme.hpp:42-49)."""
    _destroy = "phb_pme_destroy"

    def __init__(self, params: DeviceParams, box: L.Box, rc: float, precision: int = L.PHB_FP64):
        super().__init__()
        L.call("phb_pme_create", params.h, C.byref(box), rc, precision, C.byref(self.h))
        nx, ny, nz, order = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        alpha = C.c_double()
        L.call("phb_pme_info", self.h, C.byref(nx), C.byref(ny), C.byref(nz), C.byref(order), C.byref(alpha))
        self.nx, self.ny, self.nz, self.order, self.alpha = nx.value, ny.value, nz.value, order.value, alpha.value

    def energy(self, atoms, neigh, box):
        comps = L.PmeComponents()
        e = C.c_double(0)
        L.call("phb_pme_energy", self.h, atoms.h, neigh.h, C.byref(box), C.byref(comps), C.byref(e))
        return e.value, comps

    def plane_mode(self, mode: int = -1) -> int:
        """-1: report; 0: one workgroup per xy plane; 1: row and column passes.
        Returns 1 when the row and column passes are in use."""
        s = C.c_int()
        L.call("phb_pme_plane_mode", self.h, mode, C.byref(s))
        return s.value

    def spread(self, atoms):
        g = np.zeros(self.nx * self.ny * self.nz)
        L.call("phb_pme_spread", self.h, atoms.h, L.dptr(g))
        return g

    def fft_r2c(self, grid):
        grid = np.ascontiguousarray(grid, np.float64)
        k = np.zeros(2 * (self.nx // 2 + 1) * self.ny * self.nz)
        L.call("phb_pme_fft_r2c", self.h, L.dptr(grid), L.dptr(k))
        return k.view(np.complex128)

    def influence(self):
        e = np.zeros((self.nx // 2 + 1) * self.ny * self.nz)
        L.call("phb_pme_influence", self.h, L.dptr(e))
        return e

    def potential(self, grid):
        grid = np.ascontiguousarray(grid, np.float64)
        phi = np.zeros(self.nx * self.ny * self.nz)
        e = C.c_double()
        L.call("phb_pme_potential", self.h, L.dptr(grid), L.dptr(phi), C.byref(e))
        return phi, e.value


def total_energy(atoms, params, neigh, pme, box):
    """MC/src/mc.cu:43-60; returns (E, [lj, real|cut, recip, self, qcorr, slab])."""
    comps = np.zeros(6)
    e = C.c_double(0)
    L.call("phb_total_energy", atoms.h, params.h, neigh.h, pme.h if pme else None, C.byref(box), L.dptr(comps),
           C.byref(e))
    return e.value, comps


class MonteCarlo(_Handle):
    """Device-resident GC protonation / ion-swap chain (phb_mc_*).

    Replaces the trial loop of phase_gc_mu_control (MC/src/mc.cu:371-503) at a
    fixed mu_eff; the atoms handle is updated in place by accepted moves."""
    _destroy = "phb_mc_destroy"

    def __init__(self, atoms, params, neigh, pme, box, mc: McParams, seed: int, *, pending_max=8,
                 recompute_interval=1000, record_cap=0, precision=L.PHB_FP64, spec_window=8,
                 refresh=L.PHB_REFRESH_AUTO):
        super().__init__()
        self._keep = (atoms, params, neigh, pme)
        self.record_cap = record_cap
        d = L.McDesc(mc.nh_n_type, mc.nh_p_type, mc.w_type, mc.cl_type, mc.beta, mc.mu_eff, mc.w_z_extra,
                     precision, pending_max, recompute_interval, record_cap, spec_window, refresh)
        # coulombtype=coul_cut: no plan (the chain makes its own alpha-0 plan)
        L.call("phb_mc_create", atoms.h, params.h, neigh.h, pme.h if pme is not None else None, C.byref(box),
               C.byref(d), seed & 0xFFFFFFFF, C.byref(self.h))

    def run(self, ntrials: int) -> dict:
        st = L.McStats()
        L.call("phb_mc_run", self.h, ntrials, C.byref(st))
        return st.as_dict()

    def stats(self) -> dict:
        st = L.McStats()
        L.call("phb_mc_stats_get", self.h, C.byref(st))
        return st.as_dict()

    def record(self, n: int) -> dict:
        out = {"fwd": np.zeros(n, np.int8), "i": np.zeros(n, np.int32), "j": np.zeros(n, np.int32),
               "acc": np.zeros(n, np.int8), "dE": np.zeros(n), "u": np.zeros(n)}
        L.call("phb_mc_record", self.h, n, out["fwd"].ctypes.data_as(C.POINTER(C.c_byte)), L.iptr(out["i"]),
               L.iptr(out["j"]), out["acc"].ctypes.data_as(C.POINTER(C.c_byte)), L.dptr(out["dE"]), L.dptr(out["u"]))
        return out

    def delta_e(self, i, j, fwd) -> np.ndarray:
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        f = np.ascontiguousarray(fwd, np.int32)
        out = np.zeros(len(i))
        L.call("phb_mc_delta_e", self.h, len(i), L.iptr(i), L.iptr(j), L.iptr(f), L.dptr(out))
        return out

    def recompute(self) -> float:
        e = C.c_double()
        L.call("phb_mc_recompute", self.h, C.byref(e))
        return e.value

    def set_profiling(self, on: bool = True, phase_clock: bool = True):
        """HIP-event times of every launch; phase_clock also runs the delayed-acceptance
        chain's in-kernel phase clocks (which slow it slightly)."""
        L.call("phb_mc_set_profiling", self.h, 0 if not on else (1 if phase_clock else 2))

    def rng_state(self) -> np.ndarray:
        st = (C.c_uint * 625)()
        L.call("phb_mc_rng_state", self.h, st)
        return np.array(st[:], dtype=np.uint32)

    def set_refresh_tile(self, ty: int, tz: int) -> None:
        """Tile of the direct potential refresh: ty rows (8 or 4) by tz planes (4 or 2)."""
        L.call("phb_mc_set_refresh_tile", self.h, int(ty), int(tz))

    def set_mu(self, mu_eff: float) -> None:
        """Fixed mu_eff of the next runs (a titration moving one chain from pH to pH)."""
        L.call("phb_mc_set_mu", self.h, float(mu_eff))

    def set_rng_state(self, state) -> None:
        """Continue from an engine state (625 words, std::mt19937's stream form)."""
        st = np.ascontiguousarray(state, np.uint32)
        if st.shape != (625,):
            raise ValueError("engine state must hold 625 words")
        L.call("phb_mc_set_rng_state", self.h, st.ctypes.data_as(C.POINTER(C.c_uint)))

    # ---- phase 1 as main() runs it (MC/src/main.cu:180-197)
    def phase1_run(self, target_pct: float, direct_pct_threshold: int = 90, gc_block: int = 2000,
                   epoch_attempt_max: int = 100000, nacc_frac: float = 0.10, mu_eta: float = 0.3,
                   max_trials: int = 0) -> dict:
        d = L.Phase1Desc(target_pct, direct_pct_threshold, gc_block, epoch_attempt_max, nacc_frac, mu_eta, max_trials)
        r = L.Phase1Result()
        L.call("phb_phase1_run", self.h, C.byref(d), C.byref(r))
        return r.as_dict()

    def phase1_events(self) -> list:
        """What the last phase1_run did, block by block (phb_phase1_events): one dict per
        direct-flip / already-at-target / chunk / mu-update / target-reached / no-pairs event."""
        n = C.c_int()
        L.call("phb_phase1_events", self.h, None, 0, C.byref(n))
        ev = (L.Phase1Event * max(n.value, 1))()
        L.call("phb_phase1_events", self.h, ev, n.value, C.byref(n))
        return [{k: getattr(e, k) for k, _ in L.Phase1Event._fields_} for e in ev[:n.value]]

    def phase1_handover(self) -> dict:
        """The pools phase 1 handed to phase 2 (MC/include/mc.hpp:19-27)."""
        sz = np.zeros(4, np.int32)
        L.call("phb_phase1_handover", self.h, L.iptr(sz), None, 0)
        lists = np.zeros(max(int(sz.sum()), 1), np.int32)
        L.call("phb_phase1_handover", self.h, L.iptr(sz), L.iptr(lists), lists.size)
        parts = np.split(lists[:int(sz.sum())], np.cumsum(sz)[:3])
        return dict(zip(("nh_n_remaining", "nh_p_mc", "w_remaining", "cl_mc"), parts))

    # ---- phase 2 (phase2_swap_jitter's move rule, MC/src/mc.cu:159-315)
    def phase2_begin(self, patience: int, best_eps: float = 1e-9, record_cap: int = 0) -> dict:
        d = L.Phase2Desc(patience, best_eps, record_cap)
        st = L.Phase2Stats()
        L.call("phb_phase2_begin", self.h, C.byref(d), C.byref(st))
        self.p2_record_cap = record_cap
        return st.as_dict()

    def phase2_run(self, max_attempts: int) -> dict:
        st = L.Phase2Stats()
        L.call("phb_phase2_run", self.h, max_attempts, C.byref(st))
        return st.as_dict()

    def phase2_record_from(self, first: int) -> None:
        """Keep the records of trials first .. first + record_cap - 1 of phase 2."""
        L.call("phb_phase2_record_from", self.h, first)

    def phase2_record(self, n: int) -> dict:
        out = {"atoms": np.zeros((n, 4), np.int32), "acc": np.zeros(n, np.int8), "dE": np.zeros(n)}
        L.call("phb_phase2_record", self.h, n, L.iptr(out["atoms"]), out["acc"].ctypes.data_as(C.POINTER(C.c_byte)),
               L.dptr(out["dE"]))
        return out


    # ---- delayed acceptance (no reference counterpart)
    def da_attach(self, species=None, r_desc: float = 6.0, candidates: int = 64, real_space: bool = False,
                  local_pairs: bool = False, r_local: float = 0.0):
        """local_pairs: the local-pair form of the surrogate's physics part (the
        pair potential of the pool atoms within r_local (0: r_desc) or in the
        same lateral column kept exactly, the changes of the other pool atoms
        through a smoothed continuum kernel, the rest frozen by da_freeze;
        several accepted moves per launch)."""
        sp = np.ascontiguousarray(default_species(NTYPES) if species is None else species, np.int8)
        self._species = sp
        d = L.DaDesc(sp.ctypes.data_as(C.POINTER(C.c_byte)), len(sp), float(r_desc), int(candidates),
                     int(bool(real_space)), int(bool(local_pairs)), float(r_local))
        L.call("phb_da_attach", self.h, C.byref(d))

    def da_freeze(self):
        """Local-pair form: freezes the far field at the current state (call at
        the start of each new chain; the first use freezes by itself)."""
        L.call("phb_da_freeze", self.h)

    def da_descriptors(self, i, j):
        i = np.ascontiguousarray(i, np.int32); j = np.ascontiguousarray(j, np.int32)
        d = np.zeros((len(i), L.NN_IN), np.int16)
        r = np.zeros(len(i))
        L.call("phb_da_descriptors", self.h, len(i), L.iptr(i), L.iptr(j), L.sptr(d), L.dptr(r))
        return d, r

    def da_collect(self, sur: "Surrogate", i, j, fwd):
        i = np.ascontiguousarray(i, np.int32); j = np.ascontiguousarray(j, np.int32)
        f = np.ascontiguousarray(fwd, np.int32)
        L.call("phb_da_collect", self.h, sur.h, len(i), L.iptr(i), L.iptr(j), L.iptr(f))

    def da_predict(self, sur: "Surrogate", i, j, fwd) -> np.ndarray:
        i = np.ascontiguousarray(i, np.int32); j = np.ascontiguousarray(j, np.int32)
        f = np.ascontiguousarray(fwd, np.int32)
        out = np.zeros(len(i))
        L.call("phb_da_predict", self.h, sur.h, len(i), L.iptr(i), L.iptr(j), L.iptr(f), L.dptr(out))
        return out

    def da_run(self, sur: "Surrogate", ntrials: int) -> dict:
        st = L.DaStats()
        L.call("phb_da_run", self.h, sur.h, ntrials, C.byref(st))
        return st.as_dict()

    def da_record(self, n: int) -> dict:
        out = {"fwd": np.zeros(n, np.int8), "i": np.zeros(n, np.int32), "j": np.zeros(n, np.int32),
               "stage": np.zeros(n, np.int8), "dE_star": np.zeros(n), "dE_star_rev": np.zeros(n), "dE": np.zeros(n)}
        L.call("phb_da_record", self.h, n, out["fwd"].ctypes.data_as(C.POINTER(C.c_byte)), L.iptr(out["i"]),
               L.iptr(out["j"]), out["stage"].ctypes.data_as(C.POINTER(C.c_byte)), L.dptr(out["dE_star"]),
               L.dptr(out["dE_star_rev"]), L.dptr(out["dE"]))
        return out


class Surrogate(_Handle):
    """MLP surrogate of a move's Delta-E with a device dataset (phb_surrogate_*)."""
    _destroy = "phb_surrogate_destroy"

    def __init__(self, capacity: int = 0, seed: int = 1):
        super().__init__()
        L.call("phb_surrogate_create", int(capacity), seed & 0xFFFFFFFF, C.byref(self.h))

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        assert th.size == L.NN_PARAMS
        L.call("phb_surrogate_set_params", self.h, L.fptr(th))

    def get_params(self) -> np.ndarray:
        th = np.zeros(L.NN_PARAMS, np.float32)
        L.call("phb_surrogate_get_params", self.h, L.fptr(th))
        return th

    def forward(self, d) -> np.ndarray:
        d = np.ascontiguousarray(d, np.int16)
        y = np.zeros(len(d), np.float32)
        L.call("phb_surrogate_forward", self.h, len(d), L.sptr(d), L.fptr(y))
        return y

    def grad(self, d, t):
        d = np.ascontiguousarray(d, np.int16)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(L.NN_TRAIN, np.float32)
        loss = C.c_double()
        L.call("phb_surrogate_grad", self.h, len(d), L.sptr(d), L.fptr(t), L.fptr(g), C.byref(loss))
        return loss.value, g

    def add(self, d, t):
        d = np.ascontiguousarray(d, np.int16)
        t = np.ascontiguousarray(t, np.float32)
        L.call("phb_surrogate_data_add", self.h, len(d), L.sptr(d), L.fptr(t))

    def data(self, first: int = 0, n: int | None = None):
        n = self.size() - first if n is None else n
        d = np.zeros((n, L.NN_IN), np.int16)
        t = np.zeros(n, np.float32)
        L.call("phb_surrogate_data_get", self.h, first, n, L.sptr(d), L.fptr(t))
        return d, t

    def size(self) -> int:
        n = C.c_longlong()
        L.call("phb_surrogate_data_size", self.h, C.byref(n))
        return n.value

    def clear(self):
        L.call("phb_surrogate_data_clear", self.h)

    def train(self, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 1, n_train: int = 0,
              normalize: bool = True) -> dict:
        d = L.TrainDesc(epochs, batch, lr, seed & 0xFFFFFFFF, n_train, int(normalize))
        st = L.TrainStats()
        L.call("phb_surrogate_train", self.h, C.byref(d), C.byref(st))
        return st.as_dict()

    def loss_history(self, n: int) -> np.ndarray:
        out = np.zeros(n, np.float32)
        L.call("phb_surrogate_loss_history", self.h, n, L.fptr(out))
        return out

    def eval(self, first: int, n: int):
        r, a = C.c_double(), C.c_double()
        L.call("phb_surrogate_eval", self.h, first, n, C.byref(r), C.byref(a))
        return r.value, a.value


class Batch(_Handle):
    """Replicas advanced together (phb_batch_*): one workgroup per replica per
    chain launch, batched refreshes, launch pairs under a hipGraph; or one
    persistent launch per sweep (persistent=True)."""
    _destroy = "phb_batch_destroy"

    def __init__(self, mcs, graph_pairs: int = 4, use_graph: bool = True, persistent: bool = False):
        super().__init__()
        self._keep = list(mcs)
        arr = (C.c_void_p * len(mcs))(*[m.h.value for m in mcs])
        d = L.BatchDesc(graph_pairs, int(use_graph), 0, 0, 0, int(persistent))
        L.call("phb_batch_create", arr, len(mcs), C.byref(d), C.byref(self.h))

    def mode(self, persistent: bool):
        L.call("phb_batch_mode", self.h, int(bool(persistent)))

    def replica_times(self):
        """Last persistent sweep, per replica: (run ms to its finish, ms of it on refreshes)."""
        n = len(self._keep)
        run, ref = np.zeros(n), np.zeros(n)
        L.call("phb_batch_replica_times", self.h, L.dptr(run), L.dptr(ref))
        return run, ref

    def run(self, sweeps: int) -> dict:
        st = L.BatchStats()
        L.call("phb_batch_run", self.h, sweeps, C.byref(st))
        return st.as_dict()

    def autotune(self, sweeps_per_candidate: int = 1) -> dict:
        st = L.BatchStats()
        L.call("phb_batch_autotune", self.h, sweeps_per_candidate, C.byref(st))
        return st.as_dict()

    def tune(self, spec_window: int, pending_max: int, graph_pairs: int, launch_cap: int = 0):
        L.call("phb_batch_tune", self.h, spec_window, pending_max, graph_pairs, launch_cap)

    def set_refresh_tile(self, ty: int, tz: int):
        """Direct-refresh tile of every replica: ty rows (8 or 4) by tz planes (4 or 2)."""
        L.call("phb_batch_set_refresh_tile", self.h, ty, tz)


class Replica:
    """One system set up on the device the way main() does (MC/src/main.cu:106-138)."""

    def __init__(self, s: System, p: Params, mc: McParams | None = None, seed: int = 1, plane_mode: int = -1,
                 **mc_kw):
        self.s, self.p = s, p
        self.bonds_host = bond_adjacency(s.natoms, s.bonds)
        self.atoms = DeviceAtoms(s)
        self.bonds = DeviceCSR.from_host(*self.bonds_host)
        self.box = box_of(s)
        self.neigh = build_half_neighbors_gpu(self.atoms, self.box, self.bonds, p.rc)
        self.params = DeviceParams(p)
        self.pme = PMEPlan(self.params, self.box, p.rc) if p.coulombtype == L.PHB_COUL_PME else None
        if self.pme is not None and plane_mode >= 0:
            self.pme.plane_mode(plane_mode)
        self.mc = MonteCarlo(self.atoms, self.params, self.neigh, self.pme, self.box, mc, seed, **mc_kw) if mc else None

    def total_energy(self):
        return total_energy(self.atoms, self.params, self.neigh, self.pme, self.box)


def rng_seed_state(seed: int) -> np.ndarray:
    """std::mt19937(seed) in its stream form (624 words, then the index)."""
    st = np.zeros(625, np.uint32)
    L.call("phb_rng_seed_state", int(seed) & 0xFFFFFFFF, st.ctypes.data_as(C.POINTER(C.c_uint)))
    return st


def rng_draws(state, his):
    """The chains' device engine from `state`: his[k] >= 0 draws
    uniform_int_distribution<int>(0, his[k]), his[k] < 0 a uniform double (its
    bits).  Returns (draws as uint64, engine state after, serially redrawn rounds)."""
    st = np.ascontiguousarray(state, np.uint32)
    his = np.ascontiguousarray(his, np.int32)
    out = np.zeros(len(his), np.uint64)
    after = np.zeros(625, np.uint32)
    nser = C.c_longlong()
    L.call("phb_rng_draws", st.ctypes.data_as(C.POINTER(C.c_uint)), len(his), his.ctypes.data_as(C.POINTER(C.c_int)),
           out.ctypes.data_as(C.POINTER(C.c_ulonglong)), after.ctypes.data_as(C.POINTER(C.c_uint)), C.byref(nser))
    return out, after, nser.value
