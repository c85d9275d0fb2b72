"""Skin (Verlet) list (phb_verlet_*): the rc list filtered from an rc + skin
list built at a position snapshot.

Bars (2,640-atom brush, rc = 3.5, skin = 0.3, 1-2 bonds excluded):
  - right after the build, half() is bit-identical to phb_nlist_build_half(rc)
    and to the oracle's brute-force list;
  - after every atom moves by up to 0.1 (< skin / 2) the update keeps the outer
    list (no rebuild) and half() is bit-identical to a fresh rc list and the
    oracle's brute force at the new positions (no pair missing, none extra);
  - a move of 0.2 (> skin / 2) triggers exactly one rebuild, after which half()
    again equals both; the reported displacement is the largest one.
"""
import dataclasses

import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu


def _moved(s, dx, dy, dz):
    return dataclasses.replace(s, x=S.wrap_positions(s.x + dx, s.Lx), y=S.wrap_positions(s.y + dy, s.Ly),
                               z=S.wrap_positions(s.z + dz, s.Lz))


def _check(s, p, atoms, box, bonds, bh, bl, v):
    h, l = v.half(atoms).download()
    hf, lf = S.build_half_neighbors_gpu(atoms, box, bonds, p.rc).download()
    ho, lo = O.Oracle(s, p, (bh, bl)).nlist()
    assert np.array_equal(h, hf) and np.array_equal(l, lf)
    assert np.array_equal(h, ho) and np.array_equal(l, lo)
    return len(l)


def test_skin_list_equals_exact_list_until_rebuild(gpu):
    s, p, _ = S.config_small_fp64()
    bh, bl = S.bond_adjacency(s.natoms, s.bonds)
    atoms = S.DeviceAtoms(s)
    bonds = S.DeviceCSR.from_host(bh, bl)
    box = S.box_of(s)
    skin = 0.3
    v = S.VerletList(atoms, box, bonds, p.rc, skin)
    nnz0 = _check(s, p, atoms, box, bonds, bh, bl, v)
    st = v.stats()
    assert st["builds"] == 1 and st["outer_nnz"] > nnz0

    rng = np.random.default_rng(3)
    d = rng.normal(size=(3, s.natoms))
    d *= rng.uniform(0.0, 0.1, s.natoms) / np.linalg.norm(d, axis=0)
    s1 = _moved(s, *d)
    atoms.set_positions(s1.x, s1.y, s1.z)
    rebuilt, disp = v.update(atoms)
    assert not rebuilt and 0.0 < disp <= 0.1 + 1e-12
    _check(s1, p, atoms, box, bonds, bh, bl, v)
    assert v.stats()["builds"] == 1

    d2 = np.zeros((3, s.natoms))
    d2[:, 17] = 0.2 * d[:, 17] / np.linalg.norm(d[:, 17])   # along its first move: 0.2 + |d_17| in all
    s2 = _moved(s1, *d2)
    atoms.set_positions(s2.x, s2.y, s2.z)
    rebuilt, disp = v.update(atoms)
    assert rebuilt and disp >= 0.2 - 1e-12
    assert v.stats()["builds"] == 2
    _check(s2, p, atoms, box, bonds, bh, bl, v)
    rebuilt, disp = v.update(atoms)
    assert not rebuilt and disp == 0.0
