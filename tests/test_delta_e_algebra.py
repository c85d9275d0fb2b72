"""CPU check of the reciprocal-space Delta-E method the device chain uses.

The chain evaluates the SPME reciprocal change of a move from the quadratic
form E = 0.5 Q^T G Q on the charge mesh:
    dE = sum_a dq_a W_a.phi + sum_{a,b} w_ab W_a^T G W_b,   phi = G Q,
with G applied through the mesh kernel g = IFFT(eta) and separable stencil
correlations, and with moves accepted since the last refresh of phi entering
as extra pairs.  This test restates that method in numpy (on the oracle's FFT
and influence function) and checks it against the oracle's full recompute of
E_recip before and after, for single moves and for a move evaluated against a
stale phi plus one pending move.
"""
import numpy as np

from oracle import pyoracle as O
from ph_brush_amd import system as S


def _stencil(s, p, a):
    nx, ny, nz = p.grid
    P = p.pme_order
    off = P - 1 - P // 2
    first, ws = [], []
    for v, L, n in ((s.x[a], s.Lx, nx), (s.y[a], s.Ly, ny), (s.z[a], s.Lz, nz)):
        g = v / L * n
        i0 = int(np.floor(g))
        first.append((i0 - off) % n)
        ws.append(O.bspline(P, g - i0))
    return first, ws


def _w_dot(phi, s, p, a):
    nx, ny, nz = p.grid
    P = p.pme_order
    f, w = _stencil(s, p, a)
    ix = (f[0] + np.arange(P)) % nx
    iy = (f[1] + np.arange(P)) % ny
    iz = (f[2] + np.arange(P)) % nz
    blk = phi.reshape(nz, ny, nx)[np.ix_(iz, iy, ix)]
    return float(np.einsum("z,y,x,zyx->", w[2], w[1], w[0], blk))


def _cross(gk, s, p, a, b):
    nx, ny, nz = p.grid
    P = p.pme_order
    fa, wa = _stencil(s, p, a)
    fb, wb = _stencil(s, p, b)
    Cs, idx = [], []
    for ax, n in enumerate((nx, ny, nz)):
        C = np.array([sum(wa[ax][u] * wb[ax][u - d] for u in range(max(0, d), min(P, P + d)))
                      for d in range(-(P - 1), P)])
        Cs.append(C)
        idx.append((fa[ax] - fb[ax] + np.arange(-(P - 1), P)) % n)
    blk = gk.reshape(nz, ny, nx)[np.ix_(idx[2], idx[1], idx[0])]
    return float(np.einsum("z,y,x,zyx->", Cs[2], Cs[1], Cs[0], blk))


def _recip(o):
    return o.pme_energy()[1][1]


def _apply(s, mc, i, j, f):
    s2 = s.copy()
    if f:
        s2.q[i], s2.type0[i], s2.q[j], s2.type0[j] = 1.0, mc.nh_p_type - 1, -1.0, mc.cl_type - 1
    else:
        s2.q[i], s2.type0[i], s2.q[j], s2.type0[j] = 0.0, mc.nh_n_type - 1, 0.0, mc.w_type - 1
    return s2


def _moves(s, mc):
    t = s.type0 + 1
    nhn = np.where(t == mc.nh_n_type)[0]; w = np.where(t == mc.w_type)[0]
    nhp = np.where(t == mc.nh_p_type)[0]; cl = np.where(t == mc.cl_type)[0]
    return [(int(nhn[0]), int(w[-1]), 1), (int(nhp[1]), int(cl[0]), 0), (int(nhn[-1]), int(w[0]), 1)]


def test_quadratic_form_delta_e():
    for order, grid in ((4, (16, 16, 48)), (5, (16, 12, 40))):
        s, p, mc = S.config_tiny(order=order, grid=grid, n_perm=2)
        nx, ny, nz = grid
        eta = O.Oracle(s, p).influence()
        gk = O.fft_c2r_3d(eta.astype(np.complex128), nx, ny, nz)
        o0 = O.Oracle(s, p)
        o0.nlist()
        Q = o0.spread()
        phi = O.fft_c2r_3d(O.fft_r2c_3d(Q, nx, ny, nz) * eta, nx, ny, nz)
        e0 = _recip(o0)
        assert abs(0.5 * Q @ phi - e0) < 1e-10 * abs(e0)
        for (i, j, f) in _moves(s, mc):
            s1 = _apply(s, mc, i, j, f)
            dqi, dqj = s1.q[i] - s.q[i], s1.q[j] - s.q[j]
            d = (dqi * _w_dot(phi, s, p, i) + dqj * _w_dot(phi, s, p, j)
                 + 0.5 * dqi * dqi * _cross(gk, s, p, i, i) + 0.5 * dqj * dqj * _cross(gk, s, p, j, j)
                 + dqi * dqj * _cross(gk, s, p, i, j))
            o1 = O.Oracle(s1, p); o1.nlist()
            exp = _recip(o1) - e0
            assert abs(d - exp) < 1e-9 * max(1.0, abs(exp)), (order, i, j, d, exp)
        # a second move against the stale phi, with the first as a pending change
        (i1, j1, f1), (i2, j2, f2) = _moves(s, mc)[0], _moves(s, mc)[2]
        s1 = _apply(s, mc, i1, j1, f1)
        s2 = _apply(s1, mc, i2, j2, f2)
        pend = [(i1, s1.q[i1] - s.q[i1]), (j1, s1.q[j1] - s.q[j1])]
        dqi, dqj = s2.q[i2] - s1.q[i2], s2.q[j2] - s1.q[j2]
        d = (dqi * _w_dot(phi, s, p, i2) + dqj * _w_dot(phi, s, p, j2)
             + 0.5 * dqi * dqi * _cross(gk, s, p, i2, i2) + 0.5 * dqj * dqj * _cross(gk, s, p, j2, j2)
             + dqi * dqj * _cross(gk, s, p, i2, j2)
             + sum(dq * (dqi * _cross(gk, s, p, i2, a) + dqj * _cross(gk, s, p, j2, a)) for a, dq in pend))
        oa = O.Oracle(s1, p); oa.nlist()
        ob = O.Oracle(s2, p); ob.nlist()
        exp = _recip(ob) - _recip(oa)
        assert abs(d - exp) < 1e-9 * max(1.0, abs(exp)), (order, d, exp)
