"""Pins the oracle's adjoint (no reference counterpart) on the CPU:

* every transposed piece passes a dot-product test against its forward
  piece (linear pieces exactly, nonlinear ones against a central-difference
  JVP);
* the whole multi-step gradient matches central finite differences of the
  oracle's forward rollout;
* the mirror weights W make W A symmetric on grids without obstacles (and
  not with them), which is what makes the weighted CG exact.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import les_workloads as wl
from cpu_checkers import (Oracle, OrcMesh, DP, U8P, _p, _p7, oracle_rollout, oracle_adjoint,
                          PC_JACOBI, PC_CHEBYSHEV)


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.set_threads(1)
    L = o.lib
    L.orc_adj_velocity_bcs.argtypes = [DP, DP, DP, U8P, C.POINTER(OrcMesh)]
    L.orc_adj_correct.argtypes = [DP, DP, DP, DP, C.POINTER(OrcMesh), C.c_double, C.c_double]
    L.orc_adj_divergence.argtypes = [DP, DP, DP, DP, U8P, C.POINTER(OrcMesh), C.c_double,
                                     C.c_double]
    L.orc_adj_nut_bcs.argtypes = [DP, DP, U8P, C.POINTER(OrcMesh)]
    L.orc_adj_momentum.argtypes = [DP] * 11 + [C.POINTER(OrcMesh), C.c_double, C.c_double]
    L.orc_adj_sgs.argtypes = [DP] * 7 + [C.POINTER(OrcMesh), C.c_double]
    L.orc_mirror_weights.argtypes = [DP, C.POINTER(DP), C.POINTER(OrcMesh)]
    return o


M = dict(nx=9, ny=8, nz=7, dx=1 / 9, dy=0.7 / 8, dz=1.3 / 7)
N = 9 * 8 * 7


def interior_mask():
    i, j, k = np.indices((9, 8, 7)).reshape(3, -1)
    return ((i >= 1) & (i <= 7) & (j >= 1) & (j <= 6) & (k >= 1) & (k <= 5)).astype(float)


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(1)
    solid = wl.box_mask(9, 8, 7, [((3, 5), (3, 5), (2, 5))])
    solid[3 * 56 + 7 * 7 + 6] = 1  # an interior solid at i = nx-2 (outlet-copy path)
    return solid, rng


def dots(a, b):
    return sum(float(np.dot(x, y)) for x, y in zip(a, b))


def test_pieces_dot_products(orc, data):
    solid, rng = data
    R = lambda: rng.standard_normal(N)  # noqa: E731
    Mp = orc._m(M)
    L = orc.lib
    inte = interior_mask()
    d = [R(), R(), R()]
    y = [R(), R(), R()]
    # velocity BCs (affine): transpose of the linear part
    a = orc.velocity_bcs(*d, solid, M)
    z = orc.velocity_bcs(*[0 * x for x in d], solid, M)
    yb = [x.copy() for x in y]
    L.orc_adj_velocity_bcs(_p(yb[0]), _p(yb[1]), _p(yb[2]), _p(solid), Mp)
    assert np.isclose(dots([p - q for p, q in zip(a, z)], y), dots(d, yb), rtol=1e-12)
    # correction, as a function of p
    p = R()
    zero = [np.zeros(N)] * 3
    cp = [q1 - q0 for q1, q0 in zip(orc.correct(*zero, p, M, 1e-3, 1.2),
                                    orc.correct(*zero, 0 * p, M, 1e-3, 1.2))]
    pb = np.zeros(N)
    L.orc_adj_correct(_p(pb), _p(y[0]), _p(y[1]), _p(y[2]), Mp, 1e-3, 1.2)
    assert np.isclose(dots(cp, y), np.dot(p, pb), rtol=1e-12)
    # divergence -> b
    bb = R()
    b = orc.build_b(orc.divergence(*d, solid, M, 1e-3, 1.2), solid, M)
    g = [np.zeros(N) for _ in range(3)]
    L.orc_adj_divergence(_p(g[0]), _p(g[1]), _p(g[2]), _p(bb), _p(solid), Mp, 1e-3, 1.2)
    assert np.isclose(np.dot(b, bb), dots(d, g), rtol=1e-12)
    # nu_t BCs on raw interior values
    x = R() * inte
    fb = orc.nut_bcs(x, solid, M)
    yy = R()
    br = np.zeros(N)
    L.orc_adj_nut_bcs(_p(br), _p(yy), _p(solid), Mp)
    assert np.isclose(np.dot(fb, yy), np.dot(x, br), rtol=1e-12)
    # momentum (nonlinear): central-difference JVP
    u0 = [R() for _ in range(3)]
    nt = np.abs(R()) * 1e-2 * inte
    dn = R() * inte * 1e-2

    def mom(u, v, w, nt_):
        r = orc.momentum(u, v, w, np.zeros(N), np.zeros(N), np.zeros(N), nt_, M, 1e-3, 1e-3)
        return [q * inte for q in r[:3]]

    eps = 1e-6
    mp = mom(*[q + eps * e for q, e in zip(u0, d)], nt + eps * dn)
    mm = mom(*[q - eps * e for q, e in zip(u0, d)], nt - eps * dn)
    jvp = [(q1 - q0) / (2 * eps) for q1, q0 in zip(mp, mm)]
    mb = [q * inte for q in y]
    g = [np.zeros(N) for _ in range(3)]
    gn = np.zeros(N)
    L.orc_adj_momentum(*[_p(q) for q in g], _p(gn), *[_p(q) for q in mb], *[_p(q) for q in u0],
                       _p(nt), Mp, 1e-3, 1e-3)
    assert np.isclose(dots(jvp, mb), dots(d, g) + np.dot(dn, gn), rtol=1e-8)
    # SGS (nonlinear)
    def sgs(u, v, w):
        return orc.sgs(u, v, w, M, 0.17) * inte

    jv = (sgs(*[q + eps * e for q, e in zip(u0, d)]) - sgs(*[q - eps * e for q, e in zip(u0, d)]))
    jv /= 2 * eps
    yb = R() * inte
    g = [np.zeros(N) for _ in range(3)]
    L.orc_adj_sgs(*[_p(q) for q in g], _p(yb), *[_p(q) for q in u0], Mp, 0.17)
    assert np.isclose(np.dot(jv, yb), dots(d, g), rtol=1e-7)


def to_csr(A, m):
    nx, ny, nz = m["nx"], m["ny"], m["nz"]
    n = nx * ny * nz
    sij = ny * nz
    idx = np.arange(n)
    i, j, k = idx // sij, (idx % sij) // nz, idx % nz
    rows, cols, vals = [idx], [idx], [A[0]]
    for q, off, ok in ((4, -sij, i > 0), (3, sij, i < nx - 1), (2, -nz, j > 0),
                       (1, nz, j < ny - 1), (6, -1, k > 0), (5, 1, k < nz - 1)):
        sel = ok & (A[q] != 0)
        rows.append(idx[sel]); cols.append(idx[sel] + off); vals.append(A[q][sel])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


@pytest.mark.parametrize("boxes,symmetric", [([], True), ([((3, 5), (3, 5), (2, 5))], False)])
def test_mirror_weights_symmetrise(orc, boxes, symmetric):
    solid = wl.box_mask(9, 8, 7, boxes)
    A = orc.build_stencil(solid, M)
    W = np.zeros(N)
    orc.lib.orc_mirror_weights(_p(W), _p7(A), orc._m(M))
    S = sp.diags(W) @ to_csr(A, M)
    asym = abs(S - S.T).max()
    assert (asym == 0.0) == symmetric
    assert abs(to_csr(A, M) - to_csr(A, M).T).max() > 0  # A itself is not symmetric


@pytest.mark.parametrize("precond", [PC_JACOBI, PC_CHEBYSHEV])
def test_gradient_vs_finite_differences(orc, precond):
    n = 10
    m = wl.mesh(n)
    Nn = n ** 3
    solid = np.zeros(Nn, dtype=np.uint8)
    u0, v0, w0 = (wl.uniform(3, s, Nn) * 0.5 for s in range(3))
    u0, v0, w0 = orc.velocity_bcs(u0, v0, w0, solid, m)
    A = orc.build_stencil(solid, m)
    kw = dict(precond=precond, lmin=2.0, lmax=1.05 * orc.lambda_max(A, m, 40), degree=4)
    ns = 3
    box = np.zeros((n, n, n))
    box[3:7, 2:6, 3:7] = 1
    box = box.reshape(-1)
    dv = m["dx"] ** 3

    def loss(u, v, w):
        return 0.5 * np.sum(box * (u * u + v * v + w * w)) * dv

    uN, vN, wN, _, traj = oracle_rollout(orc, u0, v0, w0, solid, m, A, ns, record=True, **kw)
    gu, gv, gw = oracle_adjoint(orc, box * uN * dv, box * vN * dv, box * wN * dv, traj, solid, m,
                                A, ns, **kw)
    rng = np.random.default_rng(0)
    eps = 1e-6
    for _ in range(4):
        d = [rng.standard_normal(Nn) for _ in range(3)]
        lp = loss(*oracle_rollout(orc, u0 + eps * d[0], v0 + eps * d[1], w0 + eps * d[2], solid,
                                  m, A, ns, **kw)[:3])
        lm = loss(*oracle_rollout(orc, u0 - eps * d[0], v0 - eps * d[1], w0 - eps * d[2], solid,
                                  m, A, ns, **kw)[:3])
        fd = (lp - lm) / (2 * eps)
        ad = np.dot(gu, d[0]) + np.dot(gv, d[1]) + np.dot(gw, d[2])
        assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
