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
