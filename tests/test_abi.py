"""CPU checks of the C ABI: the library loads, exports every symbol that
include/phb.h declares, the Python bindings cover them, host-only entry points
work, and compute calls fail loudly without a GPU (there is no CPU path)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ph_brush_amd import _lib as L
from ph_brush_amd import system as S

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    txt = (ROOT / "include" / "phb.h").read_text()
    return sorted(set(re.findall(r"\b(phb_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (phb_[a-z0-9_]+)", out))
    decl = _declared()
    assert decl, "no declarations found"
    missing = [d for d in decl if d not in exported]
    assert not missing, missing
    for d in decl:
        assert hasattr(lib, d)
        assert d in L.SIGNATURES, f"{d} has no ctypes signature"


def test_abi_version():
    assert L.load().phb_abi_version() == 1


def test_generator_is_deterministic_and_counts():
    s1 = S.gen_brush(4, 4, 10, 8.0, 8.0, 20.0, n_perm=3, n_salt=5, seed=9, zfac=3.0)
    s2 = S.gen_brush(4, 4, 10, 8.0, 8.0, 20.0, n_perm=3, n_salt=5, seed=9, zfac=3.0)
    assert np.array_equal(s1.x, s2.x) and np.array_equal(s1.type0, s2.type0)
    n_beads, n_perm, n_titr = 160, 48, 112
    assert s1.natoms == n_beads + n_perm + n_titr + 10
    assert abs(s1.q.sum()) < 1e-12                      # neutral
    assert s1.Lz == 60.0 and s1.Lz_data == 20.0
    t = s1.type0 + 1
    assert np.sum(t == S.TYPE_PERM) == n_perm
    assert np.sum((t == S.TYPE_NH_N) | (t == S.TYPE_NH_P)) == n_titr
    assert np.sum((t == S.TYPE_W) | (t == S.TYPE_CL)) == n_titr + n_perm + 5
    # bonds: consecutive beads of each chain, length 0.97
    b = s1.bonds
    assert len(b) == 16 * 9
    d = np.stack([s1.x[b[:, 0]] - s1.x[b[:, 1]], s1.y[b[:, 0]] - s1.y[b[:, 1]], s1.z[b[:, 0]] - s1.z[b[:, 1]]])
    d[0] -= 8.0 * np.rint(d[0] / 8.0); d[1] -= 8.0 * np.rint(d[1] / 8.0)
    assert np.allclose(np.sqrt((d ** 2).sum(0)), 0.97)
    # no overlaps below 0.8
    from oracle import pyoracle as O
    o = O.Oracle(s1, S.Params(rc=0.8, grid=(8, 8, 8)))
    assert o.nlist(rc=0.7999)[1].size == 0


def test_generator_rejects_bad_spec():
    with pytest.raises(L.PhbError) as ei:
        S.gen_brush(2, 2, 5, -1.0, 4.0, 4.0)
    assert ei.value.code == 1


def test_compute_fails_loudly_without_gpu():
    if L.device_count() > 0:
        pytest.skip("a GPU is visible")
    s, p, _ = S.config_tiny()
    with pytest.raises(L.PhbError) as ei:
        S.DeviceAtoms(s)
    assert ei.value.code == 5 and "no CPU fallback" in str(ei.value)
