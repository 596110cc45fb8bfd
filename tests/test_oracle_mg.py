"""Pins of the oracle's multigrid preconditioner (no reference counterpart:
the reference's README.md:53 lists multigrid as a TODO), without running
the oracle's own code twice:

* the coarse operator against P^T A P formed with scipy.sparse from the
  seven arrays (an independent construction), including that the product
  keeps the 7-point pattern;
* one V-cycle against a restatement with sparse matrices (smoothing,
  restriction P^T, prolongation P, scaled correction) on several shapes;
* PCG with it solving to tolerance in far fewer iterations than
  PCG-Chebyshev(8) on the same operators;
* the mirror-weighted cycle (W A symmetric): its level-1 operator against
  P^T W A P, the cycle against the sparse restatement with the weighted
  restriction, its symmetry in the W inner product, and PCG in that inner
  product converging from a white-noise right-hand side where the
  Euclidean iteration stalls.
Runs on the CPU."""
import numpy as np
import pytest
import scipy.sparse as sp

import les_workloads as wl
from cases import random_case
from cpu_checkers import Oracle, PC_CHEBYSHEV

C_, N_, S_, E_, W_, T_, B_ = range(7)


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.set_threads(1)
    return o


def sparse_of(A, dims):
    """The operator of the seven arrays as a sparse matrix: row id has the
    centre and the six neighbours that exist (the reference's guards)."""
    nx, ny, nz = dims
    n = nx * ny * nz
    idx = np.arange(n).reshape(nx, ny, nz)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [A[C_].ravel()]
    for q, (di, dj, dk) in ((N_, (0, 1, 0)), (S_, (0, -1, 0)), (E_, (1, 0, 0)),
                            (W_, (-1, 0, 0)), (T_, (0, 0, 1)), (B_, (0, 0, -1))):
        i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        ok = ((i + di >= 0) & (i + di < nx) & (j + dj >= 0) & (j + dj < ny) &
              (k + dk >= 0) & (k + dk < nz))
        rows.append(idx[ok])
        cols.append(idx[(i + di)[ok], (j + dj)[ok], (k + dk)[ok]])
        vals.append(A[q].reshape(nx, ny, nz)[ok])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def prolongation(dims):
    nx, ny, nz = dims
    NX, NY, NZ = (nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    parent = ((i // 2) * NY + j // 2) * NZ + k // 2
    n = nx * ny * nz
    return sp.csr_matrix((np.ones(n), (np.arange(n), parent.ravel())), shape=(n, NX * NY * NZ))


def seven_of(M, dims):
    """The seven diagonals of a 7-point sparse matrix (0 where a neighbour
    does not exist), and the largest entry outside that pattern."""
    nx, ny, nz = dims
    pattern = sparse_of([np.ones(nx * ny * nz)] * 7, dims)
    outside = (abs(M) - abs(M).multiply(pattern != 0))
    out_max = abs(outside).max() if outside.nnz else 0.0
    Md = M.todok()
    idx = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    arrs = [np.zeros(nx * ny * nz) for _ in range(7)]
    for q, (di, dj, dk) in ((C_, (0, 0, 0)), (N_, (0, 1, 0)), (S_, (0, -1, 0)),
                            (E_, (1, 0, 0)), (W_, (-1, 0, 0)), (T_, (0, 0, 1)), (B_, (0, 0, -1))):
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    a, b, c = i + di, j + dj, k + dk
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        arrs[q][idx[i, j, k]] = Md.get((idx[i, j, k], idx[a, b, c]), 0.0)
    return arrs, out_max


CASES = [((7, 6, 5), "mask"), ((9, 4, 11), "mask"), ((8, 8, 8), "free"), ((5, 7, 2), "random"),
         ((10, 9, 36), "mask")]


def operator(orc, dims, kind, seed=3):
    m, solid, f = random_case(*dims, seed, frac=0.1)
    if kind == "random":  # a general StencilMatrix: every coefficient set
        return m, [f[q] + (4.0 if q == C_ else 0.0) for q in range(7)]
    if kind == "free":
        solid = np.zeros_like(solid)
    return m, orc.build_stencil(solid, m)


@pytest.mark.parametrize("dims,kind", CASES)
def test_coarse_operator_is_ptap(orc, dims, kind):
    m, A = operator(orc, dims, kind)
    Ac = orc.mg_coarsen(A, dims)
    P = prolongation(dims)
    G = (P.T @ sparse_of(A, dims) @ P).tocsr()
    cd = tuple((d + 1) // 2 for d in dims)
    ref, out_max = seven_of(G, cd)
    assert out_max == 0.0   # P^T A P keeps the 7-point pattern
    for q in range(7):
        assert np.allclose(Ac[q], ref[q], rtol=1e-13, atol=1e-13 * np.abs(ref[C_]).max()), q


def vcycle_sparse(As, dims_l, b, prm, l=0, W=None):
    """One V-cycle with sparse matrices: the same algorithm, independent code
    (W: the level-0 residual is restricted as P^T W r)."""
    A = As[l]
    D = A.diagonal()
    om = prm.omega

    def smooth(x, sweeps, zero):
        for s in range(sweeps):
            if s == 0 and zero:
                x = om * (b / D)
            else:
                x = x + om * ((b - A @ x) / D)
        return x

    if l == len(As) - 1:
        return smooth(np.zeros_like(b), prm.coarse_sweeps, True)
    x = smooth(np.zeros_like(b), prm.sweeps, True)
    P = prolongation(dims_l[l])
    res = b - A @ x
    if l == 0 and W is not None:
        res = W * res
    xc = vcycle_sparse(As, dims_l, P.T @ res, prm, l + 1)
    x = x + prm.scale * (P @ xc)
    return smooth(x, prm.sweeps, False)


@pytest.mark.parametrize("dims,kind", CASES + [((16, 12, 20), "mask")])
@pytest.mark.parametrize("sweeps,scale", [(2, 1.8), (1, 1.0)])
def test_vcycle_matches_sparse_restatement(orc, dims, kind, sweeps, scale):
    m, A = operator(orc, dims, kind, seed=5)
    prm = orc.mg_params(sweeps=sweeps, scale=scale)
    levels = orc.mg_levels(m, prm.coarse_max)
    As = [sparse_of(A, dims)]
    for l in range(1, len(levels)):
        P = prolongation(levels[l - 1])
        As.append((P.T @ As[-1] @ P).tocsr())
    r = wl.uniform(7, 1, int(np.prod(dims)))
    z = orc.mg_apply(r, A, m, sweeps=sweeps, scale=scale)
    zr = vcycle_sparse(As, levels, r, prm)
    assert np.allclose(z, zr, rtol=1e-11, atol=1e-11 * np.abs(zr).max())
    # a fixed linear operator: M(a r1 + r2) = a M r1 + M r2
    r2 = wl.uniform(7, 2, r.size)
    z2 = orc.mg_apply(r2, A, m, sweeps=sweeps, scale=scale)
    z12 = orc.mg_apply(3.0 * r + r2, A, m, sweeps=sweeps, scale=scale)
    assert np.allclose(z12, 3.0 * z + z2, rtol=1e-10, atol=1e-10 * np.abs(z12).max())


def test_levels_halve_until_coarse_max(orc):
    assert orc.mg_levels(dict(nx=64, ny=64, nz=64)) == [(64, 64, 64), (32, 32, 32),
                                                         (16, 16, 16), (8, 8, 8), (4, 4, 4)]
    assert orc.mg_levels(dict(nx=10, ny=9, nz=36)) == [(10, 9, 36), (5, 5, 18), (3, 3, 9),
                                                       (2, 2, 5), (1, 1, 3)]
    assert orc.mg_levels(dict(nx=3, ny=4, nz=2)) == [(3, 4, 2)]


@pytest.mark.parametrize("n,masked", [(32, False), (40, True)])
def test_pcg_mg_converges_faster_than_chebyshev(orc, n, masked):
    """PCG with one V-cycle per iteration reaches 1e-8 relative (checked
    on the true residual) in a fraction of PCG-Chebyshev(8)'s iterations,
    on the LES pressure operator of a spectral velocity field's
    divergence, built with and without a 5% solid mask."""
    m = wl.mesh(n)
    u, v, w = wl.spectral_velocity(n, 0)
    solid = wl.bernoulli_mask(n, 0.05, 0)
    A = orc.build_stencil(solid if masked else np.zeros_like(solid), m)
    b = orc.build_b(orc.divergence(u, v, w, solid, m, 1e-4, 1.2), solid, m)
    tol = 1e-8 * np.linalg.norm(b)
    x, it, tr = orc.pcg_mg(A, b, np.zeros_like(b), m, max_iter=500, tol=tol)
    assert 0 < it < 40, it
    assert np.linalg.norm(b - orc.matvec(x, A, m)) < 1.01 * tol
    lmax = orc.lambda_max(A, m, 20)
    _, it_ch, _ = orc.pcg(A, b, np.zeros_like(b), m, PC_CHEBYSHEV, lmin=3 * np.pi ** 2,
                          lmax=1.05 * lmax, degree=8, max_iter=5000, tol=tol)
    assert 2 * it < it_ch, (it, it_ch)


WCASES = [((7, 6, 5), 3), ((9, 4, 11), 4), ((16, 12, 20), 5), ((10, 9, 36), 6)]


def free_operator(orc, dims, seed):
    """The operator the shipped driver builds (A before the mask: no interior
    obstacles) and its mirror weights."""
    m, solid, _ = random_case(*dims, seed, frac=0.1)
    A = orc.build_stencil(np.zeros_like(solid), m)
    return m, A, orc.mirror_weights(A, m)


@pytest.mark.parametrize("dims,seed", WCASES)
def test_weighted_cycle(orc, dims, seed):
    m, A, W = free_operator(orc, dims, seed)
    S = sparse_of(A, dims)
    Ws = sp.diags(W)
    assert abs(Ws @ S - (Ws @ S).T).max() < 1e-12 * abs(S).max()   # W A is symmetric
    P = prolongation(dims)
    Ac = orc.mg_coarsen(A, dims, W)
    G = (P.T @ Ws @ S @ P).tocsr()
    cd = tuple((d + 1) // 2 for d in dims)
    ref, out_max = seven_of(G, cd)
    assert out_max == 0.0
    for q in range(7):
        assert np.allclose(Ac[q], ref[q], rtol=1e-13, atol=1e-13 * np.abs(ref[C_]).max()), q
    # the cycle against the sparse restatement (level 1 = P^T W A P)
    prm = orc.mg_params()
    levels = orc.mg_levels(m, prm.coarse_max)
    As = [S]
    for l in range(1, len(levels)):
        Pl = prolongation(levels[l - 1])
        As.append((Pl.T @ (Ws @ As[-1] if l == 1 else As[-1]) @ Pl).tocsr())
    r1 = wl.uniform(seed, 1, W.size)
    r2 = wl.uniform(seed, 2, W.size)
    z1 = orc.mg_apply(r1, A, m, W=W)
    assert np.allclose(z1, vcycle_sparse(As, levels, r1, prm, W=W), rtol=1e-11,
                       atol=1e-11 * np.abs(z1).max())
    # self-adjoint in the W inner product: (W r2).M r1 = (W r1).M r2
    z2 = orc.mg_apply(r2, A, m, W=W)
    lhs, rhs = np.dot(W * r2, z1), np.dot(W * r1, z2)
    assert abs(lhs - rhs) <= 1e-11 * np.linalg.norm(W * r1) * np.linalg.norm(z2)


@pytest.mark.parametrize("n", [24, 32])
def test_weighted_pcg_converges_from_white_noise(orc, n):
    """The first LES step's right-hand side from a U(-1,1) velocity field (the
    C1 kind of start), operator built before the mask: PCG in the W inner
    product with the weighted cycle reaches 1e-10 in a few dozen
    iterations; the Euclidean PCG with the unweighted cycle does not reach
    it in 300 (measured: it stalls near 1e-6)."""
    from cases import box_case
    from cpu_checkers import initial_state
    m, solid, f = box_case(n, 5)
    A = orc.build_stencil(np.zeros_like(solid), m)
    W = orc.mirror_weights(A, m)
    st0 = initial_state(f[0], f[1], f[2], n ** 3, solid, m, orc)
    b = orc.build_b(orc.divergence(st0["u"], st0["v"], st0["w"], solid, m, 1e-3, 1.2), solid, m)
    tol = 1e-10 * np.linalg.norm(b)
    x, it, _ = orc.pcg_mg(A, b, np.zeros_like(b), m, max_iter=300, tol=tol, W=W)
    assert 0 < it < 60, it
    assert np.linalg.norm(b - orc.matvec(x, A, m)) < 1.01 * tol
    _, it_e, _ = orc.pcg_mg(A, b, np.zeros_like(b), m, max_iter=300, tol=tol)
    assert it_e == 300
