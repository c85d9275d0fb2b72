"""The reference's file formats for Python callers (MC/include/io.hpp).

read_lammps_full / write_lammps_full (LAMMPS "full" data files), read_params
(params.in), wrap_positions and ewald_alpha_from_rtol, all through the C ABI
of libphb.so, so Python and C++ callers read a file the same way, bit for bit.
``load_run`` composes them as the reference's main() does before its first
energy (MC/src/main.cu:102-129) and returns this package's System, Params and
McParams for the device path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib as L
from . import system as S


@dataclass
class DataFile:
    """What read_lammps_full holds (MC/include/types.hpp:6-32): header counts,
    box lengths and the Atoms / Bonds rows (types 1-based, bond ends as 0-based
    row indices)."""
    natoms: int
    nbonds: int
    ntypes: int
    Lx: float
    Ly: float
    Lz: float
    id: np.ndarray
    mol: np.ndarray
    type: np.ndarray
    q: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    bond_id: np.ndarray
    bond_type: np.ndarray
    bond_i: np.ndarray
    bond_j: np.ndarray


def read_lammps_full(path) -> DataFile:
    """read_lammps_full (MC/src/io.cpp:79-146); errors raise PhbError with the reference's message."""
    h = C.c_void_p()
    info = L.DataInfo()
    L.call("phb_data_read", str(path).encode(), C.byref(h), C.byref(info))
    try:
        n, nb = info.natoms, info.nbond_rows
        ints = [np.zeros(n, np.int32) for _ in range(3)]
        reals = [np.zeros(n) for _ in range(4)]
        bonds = [np.zeros(max(nb, 1), np.int32) for _ in range(4)]
        L.call("phb_data_get", h, *(L.iptr(a) for a in ints), *(L.dptr(a) for a in reals),
               *(L.iptr(a) for a in bonds))
    finally:
        L.call("phb_data_destroy", h)
    return DataFile(info.natoms, info.nbonds, info.ntypes, info.box.Lx, info.box.Ly, info.box.Lz, *ints, *reals,
                    *(b[:nb] for b in bonds))


def write_lammps_full(path, d: DataFile, box=None) -> None:
    """write_lammps_full (MC/src/io.cpp:326-368); box (Lx, Ly, Lz) defaults to the data file's."""
    Lx, Ly, Lz = box if box is not None else (d.Lx, d.Ly, d.Lz)
    cols_i = [np.ascontiguousarray(a, np.int32) for a in (d.id, d.mol, d.type)]
    cols_d = [np.ascontiguousarray(a, np.float64) for a in (d.q, d.x, d.y, d.z)]
    nb = int(np.asarray(d.bond_id).shape[0])
    bcols = [np.ascontiguousarray(a if nb else np.zeros(1), np.int32)
             for a in (d.bond_id, d.bond_type, d.bond_i, d.bond_j)]
    v = L.DataView(d.natoms, d.nbonds, d.ntypes, int(cols_i[0].shape[0]), nb, *(L.iptr(a) for a in cols_i),
                   *(L.dptr(a) for a in cols_d), *(L.iptr(a) for a in bcols))
    L.call("phb_data_write", str(path).encode(), C.byref(v), C.byref(L.Box(Lx, Ly, Lz)))


def wrap_positions(d: DataFile) -> None:
    """wrap_positions (MC/src/io.cpp:303-316), in place, over the data file's box."""
    for name in ("x", "y", "z"):
        setattr(d, name, np.ascontiguousarray(getattr(d, name), np.float64))
    L.call("phb_wrap_positions", d.natoms, L.dptr(d.x), L.dptr(d.y), L.dptr(d.z), C.byref(L.Box(d.Lx, d.Ly, d.Lz)))


def ewald_alpha_from_rtol(rc: float, rtol: float) -> float:
    """Ewald alpha as read_params derives it (MC/src/io.cpp:46-77)."""
    a = C.c_double()
    L.call("phb_ewald_alpha_from_rtol", rc, rtol, C.byref(a))
    return a.value


@dataclass
class ParamsFile:
    """What read_params returns (Params, MC/include/types.hpp:34-69), plus the
    library's lj_rc (key lj_cutoff_nm) and nx, ny, nz (key pme_grid)."""
    rc: float
    ntypes: int
    lj_shift: int
    coulombtype: int
    epsilon_r: float
    ewald_rtol: float
    ewald_alpha: float
    pme_spacing: float
    pme_order: int
    pme_3dc: int
    pme_3dc_zfac: float
    nh_n_type: int
    nh_p_type: int
    w_type: int
    cl_type: int
    target_protonation_pct: float
    rng_seed: int
    energy_interval: int
    patience: int
    max_attempts: int
    beta: float
    w_z_extra_nm: float
    direct_pct_threshold: int
    mu_eta: float
    gc_block_fixed: int
    sgmc_epoch_attempt_max: int
    sgmc_nacc_frac_round: float
    phase2_wc: int
    lj_rc: float
    nx: int
    ny: int
    nz: int
    lj_eps: np.ndarray = None
    lj_sig: np.ndarray = None
    lj_shift_tbl: np.ndarray = None


def read_params(path, ntypes: int) -> ParamsFile:
    """read_params (MC/src/io.cpp:171-301)."""
    f = L.ParamsFile()
    t2 = max(ntypes, 0) ** 2
    eps, sig, shift = np.zeros(max(t2, 1)), np.zeros(max(t2, 1)), np.zeros(max(t2, 1))
    L.call("phb_params_read", str(path).encode(), ntypes, C.byref(f), L.dptr(eps), L.dptr(sig), L.dptr(shift))
    tables = ("lj_eps", "lj_sig", "lj_shift_tbl")
    vals = {k.name: getattr(f, k.name) for k in fields(ParamsFile) if k.name not in tables}
    return ParamsFile(**vals, lj_eps=eps[:t2], lj_sig=sig[:t2], lj_shift_tbl=shift[:t2])


def physics_params(p: ParamsFile) -> S.Params:
    """The device path's physics parameters from a params file."""
    if p.coulombtype not in (L.PHB_COUL_CUT, L.PHB_COUL_PME):
        raise ValueError("Unsupported coulombtype: unknown")   # MC/src/main.cu:168-171
    grid = (p.nx, p.ny, p.nz) if p.nx > 0 else None
    return S.Params(rc=p.rc, lj_rc=p.lj_rc, lj_shift=bool(p.lj_shift), epsilon_r=p.epsilon_r,
                    coulombtype=p.coulombtype, ewald_alpha=p.ewald_alpha, ewald_rtol=p.ewald_rtol,
                    pme_spacing=p.pme_spacing, pme_order=p.pme_order, grid=grid, pme_3dc=bool(p.pme_3dc),
                    ntypes=p.ntypes, lj_eps=p.lj_eps.copy(), lj_sig=p.lj_sig.copy())


def move_params(p: ParamsFile) -> S.McParams:
    """The chain's move parameters from a params file (mu_eff starts at 0)."""
    return S.McParams(nh_n_type=p.nh_n_type, nh_p_type=p.nh_p_type, w_type=p.w_type, cl_type=p.cl_type,
                      beta=p.beta, mu_eff=0.0, w_z_extra=p.w_z_extra_nm)


def load_run(data_path, params_path):
    """main()'s set-up (MC/src/main.cu:102-117): read the data file, ntypes from
    the largest type when the header has none, read params.in, wrap positions,
    extend Lz by max(2, zfac) under the slab correction.  Returns (DataFile,
    ParamsFile, System, Params, McParams)."""
    d = read_lammps_full(data_path)
    if d.ntypes <= 0:
        d.ntypes = int(d.type.max()) if d.natoms else 0
    pf = read_params(params_path, d.ntypes)
    wrap_positions(d)
    zfac = max(2.0, pf.pme_3dc_zfac) if pf.pme_3dc else 1.0
    bonds = np.stack([d.bond_i, d.bond_j], axis=1).astype(np.int32) if d.bond_i.size else np.zeros((0, 2), np.int32)
    s = S.System(d.x.copy(), d.y.copy(), d.z.copy(), d.q.copy(), (d.type - 1).astype(np.int32), bonds, d.Lx, d.Ly,
                 d.Lz * zfac, d.Lz, d.mol.copy())
    return d, pf, s, physics_params(pf), move_params(pf)


# ------------------------------------------------------------ workloads as input files
def data_file_of(s: S.System, ntypes: int = S.NTYPES) -> DataFile:
    """A System as the rows of a LAMMPS "full" data file: ids 1..N in row order,
    1-based types, bond type 1, the data-file box (Lz before the slab extension)."""
    n, nb = s.natoms, int(s.bonds.shape[0])
    mol = s.mol.astype(np.int32) if s.mol is not None else np.ones(n, np.int32)
    return DataFile(n, nb, ntypes, s.Lx, s.Ly, s.Lz_data or s.Lz, np.arange(1, n + 1, dtype=np.int32), mol,
                    (s.type0 + 1).astype(np.int32), s.q.copy(), s.x.copy(), s.y.copy(), s.z.copy(),
                    np.arange(1, nb + 1, dtype=np.int32), np.ones(nb, np.int32),
                    s.bonds[:, 0].astype(np.int32), s.bonds[:, 1].astype(np.int32))


def params_text(s: S.System, p: S.Params, mc: S.McParams, target_pct: float, rng_seed: int = 1, **keys) -> str:
    """params.in for a workload: the physics of p (explicit grid and LJ cutoff as
    the extension keys), the move types, beta and W filter of mc, the phase-1
    target and seed; `keys` adds or overrides entries."""
    zfac = s.Lz / s.Lz_data if s.Lz_data else 1.0
    entries = {
        "cutoff_nm": repr(float(p.rc)), "lj_cutoff_nm": repr(float(p.lj_rc)),
        "lj_shift": "yes" if p.lj_shift else "no",
        "coulombtype": "pme" if p.coulombtype == L.PHB_COUL_PME else "coul_cut",
        "epsilon_r": repr(float(p.epsilon_r)), "ewald_alpha": repr(float(p.ewald_alpha)),
        "pme_order": str(int(p.pme_order)), "pme_3dc": f"yes {float(zfac)!r}" if p.pme_3dc else "no",
        "nh_n_type": str(mc.nh_n_type), "nh_p_type": str(mc.nh_p_type), "w_type": str(mc.w_type),
        "cl_type": str(mc.cl_type), "beta": repr(float(mc.beta)), "w_z_extra_nm": repr(float(mc.w_z_extra)),
        "target_pct": repr(float(target_pct)),
        "rng_seed": str(rng_seed),
    }
    if p.grid:
        entries["pme_grid"] = " ".join(map(str, p.grid))
    entries.update({k: str(v) for k, v in keys.items()})
    T = p.ntypes
    eps, sig = np.asarray(p.lj_eps).reshape(T, T), np.asarray(p.lj_sig).reshape(T, T)
    lines = [f"{k} = {v}" for k, v in entries.items()]
    lines += [f"lj {a + 1} {b + 1} {float(eps[a, b])!r} {float(sig[a, b])!r}" for a in range(T) for b in range(a, T)]
    return "\n".join(lines) + "\n"


WORKLOADS = {
    "brush_2640": (S.config_small_fp64, 60.0),      # 64 chains x 20 beads, every bead titratable
    "diblock_44000": (S.config_30k_fp64, 60.0),     # 400 chains x 50 beads, beads 26-50 titratable
}


# ------------------------------------------------------------ presets in the reference's units
# The reference's inputs (sim_inputs/PDMAEMA_C4.zip, PDMAEMA_N1.zip,
# PMETAC_b_PDMAEMA_C4.zip, listed in .MISSING_LARGE_BLOBS) are not distributed.
# These presets stand in for them in the reference's units and conventions: nm,
# kJ/mol, e; Martini-style type ids NH_N = 4, NH_P = 5, W = 7, Cl = 8 with the
# reference's default beta = 1/(0.008314462 * 300) (MC/include/types.hpp:51-60);
# one cutoff for LJ and real-space Ewald (1.2 nm, shifted LJ), epsilon_r = 15,
# alpha from ewald_rtol, cubic splines, grid from pme_spacing 0.12 nm, slab
# correction with zfac 3.  Chains are grown by the reduced-unit generator and
# scaled by the bead diameter SIGMA_NM.  Types: 2 Na+, 3 PMETAC's quaternary
# ammonium (+1, not titratable), 4/5 DMAEMA's amine neutral/protonated, 7 the
# neutral water bead (the ion-swap partner), 8 Cl-; 1 and 6 take no atoms.
SIGMA_NM = 0.47
PRESET_TYPES = (3, 2, 4, 5, 7, 8)   # perm, cation, nh_n, nh_p, w, cl (1-based)

PRESETS = {
    # 10 x 10 PDMAEMA chains of 30 beads at 1.41 nm, 14.1 x 14.1 x 19.74 nm box: grid 120 x 120 x 500
    "pdmaema": dict(chains=(10, 10), spacing_nm=1.41, chain_len=30, n_perm=0, p_prot=0.10, n_salt=200,
                    box_nm=(14.1, 14.1, 19.74), target_pct=30.0, seed=11),
    # 11 x 11 PMETAC(15)-b-PDMAEMA(15) chains at 1.4636 nm, 16.1 x 16.1 x 19.74 nm: grid 135 x 135 x 500
    "pmetac_b_pdmaema": dict(chains=(11, 11), spacing_nm=16.1 / 11, chain_len=30, n_perm=15, p_prot=0.10,
                             n_salt=200, box_nm=(16.1, 16.1, 19.74), target_pct=30.0, seed=12),
}


def preset_system(name: str) -> S.System:
    """The preset's brush in nm (data-file box; Lz extended x3 for the slab correction)."""
    c = PRESETS[name]
    Lx, Ly, Lz = (v / SIGMA_NM for v in c["box_nm"])
    s = S.gen_brush(c["chains"][0], c["chains"][1], c["chain_len"], Lx, Ly, Lz, spacing=c["spacing_nm"] / SIGMA_NM,
                    n_perm=c["n_perm"], p_prot=c["p_prot"], n_salt=c["n_salt"], seed=c["seed"], zfac=3.0,
                    types=PRESET_TYPES)
    k = SIGMA_NM
    return S.System(s.x * k, s.y * k, s.z * k, s.q, s.type0, s.bonds, s.Lx * k, s.Ly * k, s.Lz * k, s.Lz_data * k,
                    s.mol)


def preset_params_text(name: str, rng_seed: int = 2025, **keys) -> str:
    """params.in of a preset (reference keys only; `keys` adds or overrides entries)."""
    c = PRESETS[name]
    entries = {"cutoff_nm": "1.2", "lj_shift": "yes", "coulombtype": "pme", "epsilon_r": "15",
               "ewald_rtol": "1e-5", "pme_spacing": "0.12", "pme_order": "4", "pme_3dc": "yes 3",
               "nh_n_type": "4", "nh_p_type": "5", "w_type": "7", "cl_type": "8",
               "target_pct": repr(float(c["target_pct"])), "rng_seed": str(rng_seed)}
    entries.update({k: str(v) for k, v in keys.items()})
    lines = [f"{k} = {v}" for k, v in entries.items()]
    for a in range(1, 9):
        for b in range(a, 9):
            eps = 5.0 if a == b == 7 else 3.5          # Martini-style levels: water-water stronger
            lines.append(f"lj {a} {b} {eps!r} {SIGMA_NM!r}")
    return "\n".join(lines) + "\n"


def write_preset(name: str, out_dir, rng_seed: int = 2025, **keys):
    """Writes <out_dir>/in.data and <out_dir>/params.in for a preset; returns their paths."""
    from pathlib import Path
    s = preset_system(name)
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    write_lammps_full(out / "in.data", data_file_of(s, 8))
    (out / "params.in").write_text(preset_params_text(name, rng_seed, **keys))
    return out / "in.data", out / "params.in"


def write_workload(name: str, out_dir, target_pct: float | None = None, rng_seed: int = 1, **keys):
    """Writes <out_dir>/in.data and <out_dir>/params.in for a named workload
    (the generated brushes standing in for the reference's undistributed inputs);
    returns their paths."""
    from pathlib import Path
    cfg, default_target = WORKLOADS[name]
    s, p, mc = cfg()
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    write_lammps_full(out / "in.data", data_file_of(s))
    (out / "params.in").write_text(params_text(s, p, mc, default_target if target_pct is None else target_pct,
                                               rng_seed, **keys))
    return out / "in.data", out / "params.in"


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="write a workload as a LAMMPS data file and params.in")
    ap.add_argument("workload", choices=sorted(WORKLOADS) + sorted(PRESETS))
    ap.add_argument("out_dir")
    ap.add_argument("--target-pct", type=float, default=None)
    ap.add_argument("--rng-seed", type=int, default=None)
    a = ap.parse_args()
    if a.workload in PRESETS:
        kw = {} if a.target_pct is None else {"target_pct": a.target_pct}
        paths = write_preset(a.workload, a.out_dir, 2025 if a.rng_seed is None else a.rng_seed, **kw)
    else:
        paths = write_workload(a.workload, a.out_dir, a.target_pct, 1 if a.rng_seed is None else a.rng_seed)
    for path in paths:
        print(path)
