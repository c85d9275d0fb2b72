"""SPME grids that the reference's pme_spacing gives on nm-scale boxes
(good_fft_dim(ceil(L / 0.12)), MC/src/pme.cu:46-51, :312-318): 135 x 135 x 405
(16.2 nm box, x3 slab), 180 x 180 x 500 (20 nm) and 192 x 192 x 576 (21.6 nm).  Their xy planes exceed one
workgroup's LDS, so the plane transform runs as row and column passes; 135 is
odd, so the reference's corner weight at kx = nx/2 (MC/src/pme.cu:234) enters the
potential grids of the chain.  The row and column passes are also forced on
small grids, and odd nx is run on a small grid with the one-workgroup planes.

Bars: forward transform within 1e-12 of numpy; potential grid within 1e-11 of
numpy with the convolution's influence function (eta, kx = nx/2 at half weight
for odd nx); total energy within 1e-10 of the oracle; chains equal to the
oracle's sequence with Delta-E within the chain bar of test_gpu_mc."""
import numpy as np
import pytest

from oracle import pyoracle as O
from ph_brush_amd import system as S

from test_gpu_mc import _compare_chains, _tol

pytestmark = pytest.mark.gpu


def _sys(L, seed=4):
    return S.gen_brush(6, 6, 12, L, L, L, n_perm=2, p_prot=0.5, n_salt=20, seed=seed, zfac=3.0)


def _conv_eta(eta, nx, ny, nz):
    e = eta.reshape(nz, ny, nx // 2 + 1).copy()
    if nx % 2:
        e[..., nx // 2] *= 0.5
    return e


def _check_fft_and_potential(pme, nx, ny, nz):
    g = np.random.default_rng(2).standard_normal((nz, ny, nx))
    k = pme.fft_r2c(g.ravel()).reshape(nz, ny, nx // 2 + 1)
    ref = np.fft.rfftn(g)
    assert np.abs(k - ref).max() <= 1e-12 * np.abs(ref).max()
    phi, _ = pme.potential(g.ravel())
    phi_ref = np.fft.irfftn(ref * _conv_eta(pme.influence(), nx, ny, nz), s=g.shape, axes=(0, 1, 2)) * g.size
    assert np.abs(phi.reshape(g.shape) - phi_ref).max() <= 1e-11 * np.abs(phi_ref).max()


@pytest.mark.parametrize("L,dims", [(16.2, (135, 135, 405)), (20.0, (180, 180, 500)), (21.6, (192, 192, 576))])
def test_spacing_grid_transform_and_energy(gpu, L, dims):
    s = _sys(L)
    p = S.Params(grid=None, pme_spacing=0.12)
    atoms, box = S.DeviceAtoms(s), S.box_of(s)
    dp = S.DeviceParams(p)
    pme = S.PMEPlan(dp, box, p.rc)
    assert (pme.nx, pme.ny, pme.nz) == dims and pme.plane_mode() == 1
    _check_fft_and_potential(pme, *dims)
    bh, bl = S.bond_adjacency(s.natoms, s.bonds)
    neigh = S.build_half_neighbors_gpu(atoms, box, S.DeviceCSR.from_host(bh, bl), p.rc)
    e, _ = S.total_energy(atoms, dp, neigh, pme, box)
    o = O.Oracle(s, S.Params(grid=dims), (bh, bl))
    o.nlist()
    eo, _ = o.total_energy()
    assert abs(e - eo) <= 1e-10 * max(1.0, abs(eo)), (e, eo)


@pytest.mark.parametrize("dims", [(16, 8, 32), (12, 10, 15), (9, 6, 25), (15, 16, 45), (64, 64, 256)])
def test_row_and_column_passes_on_small_grids(gpu, dims):
    nx, ny, nz = dims
    s = S.System(np.array([0.5]), np.array([0.5]), np.array([0.5]), np.array([0.0]), np.zeros(1, np.int32),
                 np.zeros((0, 2), np.int32), 10.0, 10.0, 10.0, 10.0)
    p = S.Params(grid=dims, pme_order=4, pme_3dc=False)
    pme = S.PMEPlan(S.DeviceParams(p), S.box_of(s), p.rc)
    assert pme.plane_mode() == 0 and pme.plane_mode(1) == 1
    _check_fft_and_potential(pme, nx, ny, nz)
    assert pme.plane_mode(0) == 0
    _check_fft_and_potential(pme, nx, ny, nz)


def _chain(s, p_dev, p_oracle, n, seed, **kw):
    mc = S.McParams.at_ph(7.5)
    r = S.Replica(s, p_dev, mc, seed=seed, pending_max=4, recompute_interval=0, record_cap=n, **kw)
    st = r.mc.run(n)
    assert st["attempted"] == n and st["status"] == 0 and st["accepted"] > 0
    rg = r.mc.record(n)
    o = O.Oracle(s, p_oracle, r.bonds_host)
    o.nlist()
    ro = o.gc_run(mc, n, seed)
    k = _compare_chains(rg, ro, n)
    assert k == n
    for t in range(n):
        assert abs(rg["dE"][t] - ro["dE"][t]) <= _tol(ro["dE"][t], ro["E"]), t
    q, ty = r.atoms.identity()
    assert np.array_equal(q, ro["q"]) and np.array_equal(ty, ro["type0"])
    e_re = r.mc.recompute()
    assert abs(st["E"] - e_re) <= 1e-9 * max(1.0, abs(e_re)) and abs(e_re - ro["E"]) <= 1e-9 * max(1.0, abs(ro["E"]))
    return r


def test_chain_on_the_135_spacing_grid(gpu):
    """The chain on the 16.2 nm box's own grid (row and column passes, odd nx,
    k-space refresh since nx > 64) against the oracle's full recomputes."""
    s = _sys(16.2)
    r = _chain(s, S.Params(grid=None, pme_spacing=0.12), S.Params(grid=(135, 135, 405)), 60, 3)
    assert (r.pme.nx, r.pme.plane_mode()) == (135, 1)


@pytest.mark.parametrize("plane_mode", [0, 1])
def test_chain_odd_nx_small_grid(gpu, plane_mode):
    """Odd nx with the direct refresh (nx <= 64), one-workgroup planes or forced passes."""
    s, p, _ = S.config_tiny(order=5, grid=(15, 16, 45), n_perm=2)
    _chain(s, p, p, 1500, 1, plane_mode=plane_mode)
