"""ctypes bindings of the C ABI declared in include/phb.h.

The library is loaded from ph_brush_amd/libphb.so (built in-tree by
ph_brush_amd.build).  There is no CPU fallback: if the library is missing, or
a compute call is made without a GPU, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libphb.so"

PHB_OK = 0
PHB_FP64, PHB_FP32 = 0, 1
PHB_COUL_CUT, PHB_COUL_PME, PHB_COUL_UNKNOWN = 0, 1, 2
PHB_REFRESH_AUTO, PHB_REFRESH_DIRECT, PHB_REFRESH_KSPACE = 0, 1, 2

_ERRNAMES = {1: "ARG", 2: "HIP", 3: "ALLOC", 4: "STATE", 5: "UNSUPPORTED", 6: "INTERNAL"}


class PhbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[PHB_ERR_{_ERRNAMES.get(code, code)}] {msg}")
        self.code = code


class Box(C.Structure):
    _fields_ = [("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double)]


class ParamsDesc(C.Structure):
    _fields_ = [
        ("ntypes", C.c_int), ("lj_eps", C.POINTER(C.c_double)), ("lj_sig", C.POINTER(C.c_double)),
        ("lj_shift", C.c_int), ("rc", C.c_double), ("lj_rc", C.c_double), ("epsilon_r", C.c_double),
        ("coulombtype", C.c_int), ("ewald_alpha", C.c_double), ("ewald_rtol", C.c_double),
        ("pme_spacing", C.c_double), ("pme_order", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
        ("nz", C.c_int), ("pme_3dc", C.c_int),
    ]


class PmeComponents(C.Structure):
    _fields_ = [("real_space", C.c_double), ("recip_space", C.c_double), ("self_term", C.c_double),
                ("qcorr_term", C.c_double), ("slab_term", C.c_double)]


class BrushSpec(C.Structure):
    _fields_ = [
        ("chains_x", C.c_int), ("chains_y", C.c_int), ("graft_spacing", C.c_double), ("graft_z", C.c_double),
        ("chain_len", C.c_int), ("n_perm", C.c_int), ("p_prot", C.c_double), ("n_salt", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double), ("bond", C.c_double),
        ("min_dist", C.c_double), ("seed", C.c_ulonglong),
        ("type_perm", C.c_int), ("type_cation", C.c_int), ("type_nh_n", C.c_int), ("type_nh_p", C.c_int),
        ("type_w", C.c_int), ("type_cl", C.c_int),
    ]


class McDesc(C.Structure):
    _fields_ = [
        ("nh_n_type", C.c_int), ("nh_p_type", C.c_int), ("w_type", C.c_int), ("cl_type", C.c_int),
        ("beta", C.c_double), ("mu_eff", C.c_double), ("w_z_extra", C.c_double),
        ("precision", C.c_int), ("pending_max", C.c_int), ("recompute_interval", C.c_int),
        ("record_cap", C.c_longlong), ("spec_window", C.c_int), ("refresh", C.c_int),
    ]


class McStats(C.Structure):
    _fields_ = [
        ("attempted", C.c_longlong), ("accepted", C.c_longlong), ("accepted_forward", C.c_longlong),
        ("accepted_backward", C.c_longlong), ("nprot", C.c_int), ("nsites", C.c_int), ("status", C.c_int),
        ("E", C.c_double), ("E_recomputed", C.c_double), ("max_drift", C.c_double),
        ("refreshes", C.c_longlong), ("recomputes", C.c_longlong), ("launches", C.c_longlong),
        ("chain_ms", C.c_double), ("refresh_ms", C.c_double), ("recompute_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Phase1Desc(C.Structure):
    _fields_ = [("target_pct", C.c_double), ("direct_pct_threshold", C.c_int), ("gc_block", C.c_int),
                ("epoch_attempt_max", C.c_int), ("nacc_frac", C.c_double), ("mu_eta", C.c_double),
                ("max_trials", C.c_longlong), ("on_event", C.c_void_p), ("user", C.c_void_p)]


class Phase1Result(C.Structure):
    _fields_ = [("attempted", C.c_longlong), ("accepted", C.c_longlong), ("accepted_forward", C.c_longlong),
                ("accepted_backward", C.c_longlong), ("blocks_done", C.c_int), ("Nh", C.c_int), ("Nprot", C.c_int),
                ("target_count", C.c_int), ("reached_target", C.c_int), ("direct", C.c_int), ("status", C.c_int),
                ("E", C.c_double), ("mu_eff", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Phase2Desc(C.Structure):
    _fields_ = [("patience", C.c_int), ("best_eps", C.c_double), ("record_cap", C.c_longlong)]


class Phase2Stats(C.Structure):
    _fields_ = [("attempted", C.c_longlong), ("accepted", C.c_longlong), ("mode", C.c_int), ("stop", C.c_int),
                ("since_best", C.c_int), ("status", C.c_int), ("E", C.c_double), ("best_E", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrainDesc(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_double), ("seed", C.c_uint),
                ("n_train", C.c_longlong), ("normalize", C.c_int)]


class TrainStats(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("first_loss", C.c_double), ("last_loss", C.c_double), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DaDesc(C.Structure):
    _fields_ = [("species_of_type", C.POINTER(C.c_byte)), ("ntypes", C.c_int), ("r_desc", C.c_double),
                ("candidates", C.c_int), ("real_space", C.c_int), ("local_pairs", C.c_int),
                ("r_local", C.c_double)]


class DaStats(C.Structure):
    _fields_ = [("attempted", C.c_longlong), ("stage1_accepted", C.c_longlong), ("accepted", C.c_longlong),
                ("rounds", C.c_longlong), ("launches", C.c_longlong), ("nprot", C.c_int), ("nsites", C.c_int),
                ("status", C.c_int), ("E", C.c_double), ("chain_ms", C.c_double), ("refresh_ms", C.c_double),
                ("phase_us", C.c_double * 8)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["phase_us"] = list(self.phase_us)
        return d


class BatchDesc(C.Structure):
    _fields_ = [("graph_pairs", C.c_int), ("use_graph", C.c_int), ("spec_window", C.c_int), ("pending_max", C.c_int),
                ("launch_cap", C.c_longlong), ("persistent", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("attempted", C.c_longlong), ("accepted", C.c_longlong), ("launches", C.c_longlong),
                ("replays", C.c_longlong), ("ms", C.c_double), ("trials_per_s", C.c_double),
                ("spec_window", C.c_int), ("pending_max", C.c_int), ("graph_pairs", C.c_int),
                ("launch_cap", C.c_longlong), ("persistent", C.c_int), ("refresh_ty", C.c_int),
                ("refresh_tz", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


NN_IN, NN_HIDDEN, NN_TRAIN, NN_PARAMS = 48, 128, 22913, 22963

_h = C.c_void_p
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_pll = C.POINTER(C.c_longlong)
_pf = C.POINTER(C.c_float)
_ps = C.POINTER(C.c_short)

class DataInfo(C.Structure):
    _fields_ = [("natoms", C.c_int), ("nbonds", C.c_int), ("ntypes", C.c_int), ("nbond_rows", C.c_int),
                ("box", Box)]


class DataView(C.Structure):
    _fields_ = [("natoms", C.c_int), ("nbonds", C.c_int), ("ntypes", C.c_int), ("natoms_rows", C.c_int),
                ("nbond_rows", C.c_int), ("id", _pi), ("mol", _pi), ("type", _pi), ("q", _pd), ("x", _pd),
                ("y", _pd), ("z", _pd), ("bond_id", _pi), ("bond_type", _pi), ("bond_i", _pi), ("bond_j", _pi)]


class ParamsFile(C.Structure):
    _fields_ = [("rc", C.c_double), ("ntypes", C.c_int), ("lj_shift", C.c_int), ("coulombtype", C.c_int),
                ("epsilon_r", C.c_double), ("ewald_rtol", C.c_double), ("ewald_alpha", C.c_double),
                ("pme_spacing", C.c_double), ("pme_order", C.c_int), ("pme_3dc", C.c_int),
                ("pme_3dc_zfac", C.c_double), ("nh_n_type", C.c_int), ("nh_p_type", C.c_int), ("w_type", C.c_int),
                ("cl_type", C.c_int), ("target_protonation_pct", C.c_double), ("rng_seed", C.c_int),
                ("energy_interval", C.c_int), ("patience", C.c_int), ("max_attempts", C.c_int),
                ("beta", C.c_double), ("w_z_extra_nm", C.c_double), ("direct_pct_threshold", C.c_int),
                ("mu_eta", C.c_double), ("gc_block_fixed", C.c_int), ("sgmc_epoch_attempt_max", C.c_int),
                ("sgmc_nacc_frac_round", C.c_double), ("phase2_wc", C.c_int), ("lj_rc", C.c_double),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)]


PHB_P1_DIRECT, PHB_P1_AT_TARGET, PHB_P1_CHUNK, PHB_P1_MU_UPDATE, PHB_P1_REACHED, PHB_P1_NO_PAIRS = 1, 2, 3, 4, 5, 6


class Phase1Event(C.Structure):
    _fields_ = [("kind", C.c_int), ("epoch", C.c_int), ("attempted", C.c_longlong), ("accepted", C.c_longlong),
                ("fwd_total", C.c_longlong), ("bwd_total", C.c_longlong), ("acc_since", C.c_longlong),
                ("att_since", C.c_longlong), ("block_attempted", C.c_longlong), ("block_accepted", C.c_longlong),
                ("step_in_block", C.c_int), ("early_stop", C.c_int), ("nprot", C.c_int), ("nh", C.c_int),
                ("target_count", C.c_int), ("remaining", C.c_int), ("r0", C.c_int), ("nacc_target", C.c_int),
                ("amax", C.c_int), ("gc_block", C.c_int), ("mu_over_kT", C.c_double), ("E", C.c_double),
                ("f", C.c_double), ("w_info", C.c_double)]


# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "phb_last_error": [],
    "phb_abi_version": [],
    "phb_device_count": [_pi],
    "phb_atoms_create": [C.c_int, _pd, _pd, _pd, _pd, _pi, C.POINTER(_h)],
    "phb_atoms_destroy": [_h],
    "phb_atoms_natoms": [_h, _pi],
    "phb_atoms_set_identity": [_h, C.c_int, _pi, _pi, _pd],
    "phb_atoms_get_identity": [_h, _pd, _pi],
    "phb_csr_create": [C.c_int, _pi, _pi, C.POINTER(_h)],
    "phb_csr_destroy": [_h],
    "phb_csr_shape": [_h, _pi, _pll],
    "phb_csr_download": [_h, _pi, _pi],
    "phb_nlist_build_half": [_h, C.POINTER(Box), _h, C.c_double, C.POINTER(_h)],
    "phb_atoms_set_positions": [_h, _pd, _pd, _pd],
    "phb_phase1_run": [_h, C.POINTER(Phase1Desc), C.POINTER(Phase1Result)],
    "phb_phase1_handover": [_h, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_longlong],
    "phb_phase2_set_pools": [_h, _pi, _pi],
    "phb_phase2_begin": [_h, C.POINTER(Phase2Desc), C.POINTER(Phase2Stats)],
    "phb_phase2_run": [_h, C.c_longlong, C.POINTER(Phase2Stats)],
    "phb_phase2_record_from": [_h, C.c_longlong],
    "phb_phase2_record": [_h, C.c_longlong, _pi, C.POINTER(C.c_byte), _pd],
    "phb_verlet_create": [_h, C.POINTER(Box), _h, C.c_double, C.c_double, C.POINTER(_h)],
    "phb_verlet_destroy": [_h],
    "phb_verlet_update": [_h, _h, C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "phb_verlet_half": [_h, _h, C.POINTER(_h)],
    "phb_verlet_stats": [_h, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_double)],
    "phb_params_create": [C.POINTER(ParamsDesc), C.POINTER(_h)],
    "phb_params_destroy": [_h],
    "phb_energy_lj": [_h, _h, _h, C.POINTER(Box), _pd],
    "phb_energy_coul_cut": [_h, _h, _h, C.POINTER(Box), _pd],
    "phb_pme_create": [_h, C.POINTER(Box), C.c_double, C.c_int, C.POINTER(_h)],
    "phb_pme_destroy": [_h],
    "phb_pme_plane_mode": [_h, C.c_int, _pi],
    "phb_pme_info": [_h, _pi, _pi, _pi, _pi, _pd],
    "phb_pme_energy": [_h, _h, _h, C.POINTER(Box), C.POINTER(PmeComponents), _pd],
    "phb_total_energy": [_h, _h, _h, _h, C.POINTER(Box), _pd, _pd],
    "phb_total_energy_noalloc": [_h, _h, _h, _h, C.POINTER(Box), _pd, _pd],
    "phb_pme_spread": [_h, _h, _pd],
    "phb_pme_fft_r2c": [_h, _pd, _pd],
    "phb_pme_influence": [_h, _pd],
    "phb_pme_potential": [_h, _pd, _pd, _pd],
    "phb_mc_create": [_h, _h, _h, _h, C.POINTER(Box), C.POINTER(McDesc), C.c_uint, C.POINTER(_h)],
    "phb_mc_destroy": [_h],
    "phb_mc_run": [_h, C.c_longlong, C.POINTER(McStats)],
    "phb_mc_stats_get": [_h, C.POINTER(McStats)],
    "phb_mc_record": [_h, C.c_longlong, C.POINTER(C.c_byte), _pi, _pi, C.POINTER(C.c_byte), _pd, _pd],
    "phb_mc_delta_e": [_h, C.c_int, _pi, _pi, _pi, _pd],
    "phb_mc_recompute": [_h, _pd],
    "phb_mc_rng_state": [_h, C.POINTER(C.c_uint)],
    "phb_mc_set_rng_state": [_h, C.POINTER(C.c_uint)],
    "phb_mc_set_mu": [_h, C.c_double],
    "phb_mc_set_refresh_tile": [_h, C.c_int, C.c_int],
    "phb_rng_seed_state": [C.c_uint, C.POINTER(C.c_uint)],
    "phb_rng_draws": [C.POINTER(C.c_uint), C.c_longlong, _pi, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint),
                      C.POINTER(C.c_longlong)],
    "phb_mc_set_profiling": [_h, C.c_int],
    "phb_surrogate_create": [C.c_longlong, C.c_uint, C.POINTER(_h)],
    "phb_surrogate_destroy": [_h],
    "phb_surrogate_set_params": [_h, _pf],
    "phb_surrogate_get_params": [_h, _pf],
    "phb_surrogate_forward": [_h, C.c_int, _ps, _pf],
    "phb_surrogate_grad": [_h, C.c_int, _ps, _pf, _pf, _pd],
    "phb_surrogate_data_add": [_h, C.c_int, _ps, _pf],
    "phb_surrogate_data_get": [_h, C.c_longlong, C.c_int, _ps, _pf],
    "phb_surrogate_data_size": [_h, _pll],
    "phb_surrogate_data_clear": [_h],
    "phb_surrogate_train": [_h, C.POINTER(TrainDesc), C.POINTER(TrainStats)],
    "phb_surrogate_loss_history": [_h, C.c_longlong, _pf],
    "phb_surrogate_eval": [_h, C.c_longlong, C.c_longlong, _pd, _pd],
    "phb_da_attach": [_h, C.POINTER(DaDesc)],
    "phb_da_descriptors": [_h, C.c_int, _pi, _pi, _ps, _pd],
    "phb_da_freeze": [_h],
    "phb_da_collect": [_h, _h, C.c_int, _pi, _pi, _pi],
    "phb_da_predict": [_h, _h, C.c_int, _pi, _pi, _pi, _pd],
    "phb_da_run": [_h, _h, C.c_longlong, C.POINTER(DaStats)],
    "phb_da_record": [_h, C.c_longlong, C.POINTER(C.c_byte), _pi, _pi, C.POINTER(C.c_byte), _pd, _pd, _pd],
    "phb_batch_create": [C.POINTER(_h), C.c_int, C.POINTER(BatchDesc), C.POINTER(_h)],
    "phb_batch_destroy": [_h],
    "phb_batch_run": [_h, C.c_int, C.POINTER(BatchStats)],
    "phb_batch_autotune": [_h, C.c_int, C.POINTER(BatchStats)],
    "phb_batch_tune": [_h, C.c_int, C.c_int, C.c_int, C.c_longlong],
    "phb_batch_mode": [_h, C.c_int],
    "phb_batch_set_refresh_tile": [_h, C.c_int, C.c_int],
    "phb_batch_replica_times": [_h, _pd, _pd],
    "phb_data_read": [C.c_char_p, C.POINTER(_h), C.POINTER(DataInfo)],
    "phb_data_get": [_h, _pi, _pi, _pi, _pd, _pd, _pd, _pd, _pi, _pi, _pi, _pi],
    "phb_data_destroy": [_h],
    "phb_data_write": [C.c_char_p, C.POINTER(DataView), C.POINTER(Box)],
    "phb_wrap_positions": [C.c_int, _pd, _pd, _pd, C.POINTER(Box)],
    "phb_ewald_alpha_from_rtol": [C.c_double, C.c_double, _pd],
    "phb_params_read": [C.c_char_p, C.c_int, C.POINTER(ParamsFile), _pd, _pd, _pd],
    "phb_phase1_events": [_h, C.POINTER(Phase1Event), C.c_int, _pi],
    "phb_gen_brush": [C.POINTER(BrushSpec), C.c_int, _pd, _pd, _pd, _pd, _pi, _pi, _pi, _pi, C.c_int, _pi, _pi],
}
_RESTYPE = {"phb_last_error": C.c_char_p}

_lib = None


def load() -> C.CDLL:
    """Load libphb.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ph_brush_amd.build` "
                          "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        if not hasattr(lib, name):
            continue
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call a status-returning entry point; raise PhbError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != PHB_OK:
        raise PhbError(rc, lib.phb_last_error().decode(errors="replace"))


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_pd)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_pf)


def sptr(a: np.ndarray):
    assert a.dtype == np.int16 and a.flags.c_contiguous
    return a.ctypes.data_as(_ps)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_pi)


def device_count() -> int:
    n = C.c_int(0)
    call("phb_device_count", C.byref(n))
    return n.value
