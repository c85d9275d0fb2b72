"""The scratch-free total energy (phb_total_energy_noalloc) against
phb_total_energy through the C ABI, on two calls in a row, the second on a
smaller system, for both coulombtypes; the oracle's total besides.

Bar: coul_cut components bit-identical (the same kernels and fixed-order
reductions, no atomics); PME components within 1e-12 relative (the spread's
FP64 atomics may reorder the last bit) and the total within 1e-10 of the oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import _lib as L
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu


def _noalloc(r):
    comps = np.zeros(6)
    e = C.c_double(0)
    L.call("phb_total_energy_noalloc", r.atoms.h, r.params.h, r.neigh.h, r.pme.h if r.pme else None,
           C.byref(r.box), L.dptr(comps), C.byref(e))
    return e.value, comps


def _close(a, b, rel):
    return abs(a - b) <= rel * max(1.0, abs(b))


@pytest.mark.parametrize("coul", [L.PHB_COUL_PME, L.PHB_COUL_CUT])
def test_noalloc_total_energy_equals_total_energy(gpu, coul):
    big = S.config_small_fp64()
    small = S.config_tiny(order=5, grid=(16, 16, 48), n_perm=2)
    for s, p, _ in (big, small):
        p.coulombtype = coul
        r = S.Replica(s, p)
        e0, c0 = r.total_energy()
        e1, c1 = _noalloc(r)
        if coul == L.PHB_COUL_CUT:
            assert e1 == e0 and np.array_equal(c1, c0)
        else:
            assert _close(e1, e0, 1e-12) and all(_close(a, b, 1e-12) for a, b in zip(c1, c0))
        o = O.Oracle(s, p, r.bonds_host)
        o.nlist()
        eo, _ = o.total_energy()
        assert _close(e1, eo, 1e-10), (s.natoms, e1, eo)
