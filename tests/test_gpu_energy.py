"""GPU parity of the full-energy path against the oracle (through the C ABI).

Bar: neighbour CSR bit-exact; FP64 energies within 1e-10 relative (scale =
max(|component|, 1)); FFT within 1e-12 of numpy relative to the spectrum norm.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import _lib as L
from ph_brush_amd import system as S

pytestmark = pytest.mark.gpu

REL = 1e-10


def _setup(s, p):
    bh, bl = S.bond_adjacency(s.natoms, s.bonds)
    atoms = S.DeviceAtoms(s)
    bonds = S.DeviceCSR.from_host(bh, bl)
    box = S.box_of(s)
    neigh = S.build_half_neighbors_gpu(atoms, box, bonds, p.rc)
    dp = S.DeviceParams(p)
    o = O.Oracle(s, p, (bh, bl))
    return atoms, bonds, box, neigh, dp, o


CASES = {
    "tiny_o4": lambda: S.config_tiny(order=4, grid=(16, 16, 48)),
    "tiny_o5_mixed": lambda: S.config_tiny(order=5, grid=(12, 10, 30), n_perm=2),
    "small_2640": lambda: S.config_small_fp64(),
}


@pytest.fixture(scope="module", params=list(CASES))
def case(request, gpu):
    s, p, m = CASES[request.param]()
    return (request.param, s, p) + _setup(s, p)


def test_nlist_bit_exact(case):
    name, s, p, atoms, bonds, box, neigh, dp, o = case
    h, l = neigh.download()
    ho, lo = o.nlist()
    assert np.array_equal(h, ho)
    assert np.array_equal(l, lo)


def _close(a, b, rel=REL):
    return abs(a - b) <= rel * max(abs(b), 1.0)


def test_short_range_energies(case):
    name, s, p, atoms, bonds, box, neigh, dp, o = case
    o.nlist()
    assert _close(S.compute_lj_sr_energy_gpu(atoms, dp, neigh, box), o.energy_lj())
    assert _close(S.compute_coul_sr_energy_gpu(atoms, dp, neigh, box), o.energy_coul_cut())


def test_pme_energy_components(case):
    name, s, p, atoms, bonds, box, neigh, dp, o = case
    o.nlist()
    pme = S.PMEPlan(dp, box, p.rc)
    assert (pme.nx, pme.ny, pme.nz, pme.order) == (*p.grid, p.pme_order)
    e, c = pme.energy(atoms, neigh, box)
    eo, co = o.pme_energy()   # real, recip, self, qcorr, slab
    assert _close(c.real_space, co[0])
    assert _close(c.recip_space, co[1])
    assert _close(c.self_term, co[2])
    assert _close(c.qcorr_term, co[3] + co[4])
    assert _close(c.slab_term, co[4])
    assert _close(e, eo)
    et, ct = S.total_energy(atoms, dp, neigh, pme, box)
    eto, cto = o.total_energy()
    assert _close(et, eto)
    assert np.all([_close(a, b) for a, b in zip(ct, cto)])


def test_spread_and_influence(case):
    name, s, p, atoms, bonds, box, neigh, dp, o = case
    pme = S.PMEPlan(dp, box, p.rc)
    g = pme.spread(atoms)
    go = o.spread()
    assert np.abs(g - go).max() <= 1e-13 * max(1.0, np.abs(go).max())
    ei, eo = pme.influence(), o.influence()
    assert np.abs(ei - eo).max() <= 1e-14 * np.abs(eo).max()


def test_potential_grid(case):
    name, s, p, atoms, bonds, box, neigh, dp, o = case
    pme = S.PMEPlan(dp, box, p.rc)
    nx, ny, nz = p.grid
    g = o.spread()
    phi, e = pme.potential(g)
    eta = o.influence()
    k = O.fft_r2c_3d(g, nx, ny, nz)
    phi_o = O.fft_c2r_3d(k * eta, nx, ny, nz)
    assert np.abs(phi - phi_o).max() <= 1e-11 * np.abs(phi_o).max()
    # E_recip = 0.5 * Q . phi equals the reference's half-spectrum sum
    o.nlist()
    _, co = o.pme_energy()
    assert _close(e, co[1])
    assert _close(0.5 * float(np.dot(g, phi)), co[1], 1e-9)


@pytest.mark.parametrize("dims", [(16, 8, 32), (12, 10, 15), (32, 32, 128), (64, 64, 256), (20, 18, 45), (9, 6, 25)])
def test_fft_r2c_matches_numpy(gpu, dims):
    nx, ny, nz = dims
    s = S.System(np.array([0.5]), np.array([0.5]), np.array([0.5]), np.array([0.0]), np.zeros(1, np.int32),
                 np.zeros((0, 2), np.int32), 10.0, 10.0, 10.0, 10.0)
    p = S.Params(grid=dims, pme_order=4, pme_3dc=False)
    pme = S.PMEPlan(S.DeviceParams(p), S.box_of(s), p.rc)
    g = np.random.default_rng(2).standard_normal((nz, ny, nx))
    k = pme.fft_r2c(g.ravel()).reshape(nz, ny, nx // 2 + 1)
    ref = np.fft.rfftn(g)
    assert np.abs(k - ref).max() <= 1e-12 * np.abs(ref).max()


def test_identity_roundtrip_and_errors(gpu):
    s, p, _ = S.config_tiny()
    atoms = S.DeviceAtoms(s)
    atoms.set_identity([3, 5], [S.TYPE_NH_P - 1, S.TYPE_CL - 1], [1.0, -1.0])
    q, t = atoms.identity()
    exp_q = s.q.copy(); exp_q[[3, 5]] = [1.0, -1.0]
    exp_t = s.type0.copy(); exp_t[[3, 5]] = [S.TYPE_NH_P - 1, S.TYPE_CL - 1]
    assert np.array_equal(q, exp_q) and np.array_equal(t, exp_t)
    with pytest.raises(L.PhbError) as ei:
        atoms.set_identity([s.natoms], [0], [0.0])
    assert ei.value.code == 1 and "out of range" in str(ei.value)
    with pytest.raises(L.PhbError):
        S.build_half_neighbors_gpu(atoms, S.box_of(s), None, -1.0)
    bad = S.Params(rc=0.0)
    with pytest.raises(L.PhbError) as ei:
        S.DeviceParams(bad)
    assert "cutoff_nm must be > 0" in str(ei.value)


def test_30k_energy_and_list(gpu):
    s, p, _ = S.config_30k_fp64()
    atoms, bonds, box, neigh, dp, o = _setup(s, p)
    h, l = neigh.download()
    ho, lo = o.nlist()
    assert np.array_equal(h, ho) and np.array_equal(l, lo)
    pme = S.PMEPlan(dp, box, p.rc)
    et, ct = S.total_energy(atoms, dp, neigh, pme, box)
    eto, cto = o.total_energy()
    assert _close(et, eto)
    assert np.all([_close(a, b) for a, b in zip(ct, cto)])


def _permuted(s, perm):
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    return S.System(s.x[perm], s.y[perm], s.z[perm], s.q[perm], s.type0[perm], inv[s.bonds].astype(np.int32),
                    s.Lx, s.Ly, s.Lz, s.Lz_data)


def test_spread_bit_identical_under_atom_order(gpu):
    """The tiled spread sums in fixed point, so the grid of the 44,000-atom brush
    (tiles of about 280 atoms in the brush) is the same bit for bit on a repeat
    and with the atoms in a random order, and within 1e-13 of the oracle's FP64
    sum."""
    s, p, _ = S.config_30k_fp64()
    box = S.box_of(s)
    pme = S.PMEPlan(S.DeviceParams(p), box, p.rc)
    g1 = pme.spread(S.DeviceAtoms(s))
    g2 = pme.spread(S.DeviceAtoms(s))
    perm = np.random.default_rng(11).permutation(s.natoms)
    g3 = pme.spread(S.DeviceAtoms(_permuted(s, perm)))
    assert np.array_equal(g1.view(np.int64), g2.view(np.int64))
    assert np.array_equal(g1.view(np.int64), g3.view(np.int64))
    bh, bl = S.bond_adjacency(s.natoms, s.bonds)
    go = O.Oracle(s, p, (bh, bl)).spread()
    assert np.abs(g1 - go).max() <= 1e-13 * max(1.0, np.abs(go).max())
    assert abs(g1.sum() - s.q.sum()) <= 1e-9


@pytest.mark.parametrize("order", [3, 4, 5, 6, 8])
def test_spread_crowded_tile(gpu, order):
    """700 charged atoms inside one 8^3-point tile (staged in six chunks), beside
    sparse atoms, on a grid whose edges are not multiples of the tile (narrow last
    tiles, stencils that wrap): oracle within 1e-13, bit-identical under atom
    order."""
    rng = np.random.default_rng(order)
    nx, ny, nz = 20, 18, 45
    Lx, Ly, Lz = 10.0, 9.0, 22.5                 # grid spacing 0.5
    n_dense, n_sparse = 700, 300
    # grid coordinates 4.4-7.6: stencil origins 1-6 for every order, all in the first tile
    x = np.concatenate([rng.uniform(2.2, 3.8, n_dense), rng.uniform(0, Lx, n_sparse)])
    y = np.concatenate([rng.uniform(2.2, 3.8, n_dense), rng.uniform(0, Ly, n_sparse)])
    z = np.concatenate([rng.uniform(2.2, 3.8, n_dense), rng.uniform(0, Lz, n_sparse)])
    q = rng.choice([-1.0, 0.0, 0.5, 1.0], n_dense + n_sparse)
    n = n_dense + n_sparse
    s = S.System(x, y, z, q, np.zeros(n, np.int32), np.zeros((0, 2), np.int32), Lx, Ly, Lz, Lz)
    p = S.Params(grid=(nx, ny, nz), pme_order=order, pme_3dc=False)
    pme = S.PMEPlan(S.DeviceParams(p), S.box_of(s), p.rc)
    g = pme.spread(S.DeviceAtoms(s))
    bh, bl = S.bond_adjacency(n, s.bonds)
    go = O.Oracle(s, p, (bh, bl)).spread()
    assert np.abs(g - go).max() <= 1e-13 * max(1.0, np.abs(go).max())
    perm = rng.permutation(n)
    g2 = pme.spread(S.DeviceAtoms(_permuted(s, perm)))
    assert np.array_equal(g.view(np.int64), g2.view(np.int64))
