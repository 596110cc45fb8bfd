"""Pins the oracle (oracle/les_oracle.c) to the unmodified reference solver.

The reference is compiled from its own sources (oracle/Makefile ->
oracle/_ref/libref.so).  On one thread both must agree bit for bit, function
by function and over whole steps.  Skipped where that build is absent (the
committed fixtures in tests/golden/ pin the oracle there instead).
"""
import os

import numpy as np
import pytest

import les_workloads as wl
from cpu_checkers import (Oracle, Reference, REF_SO, initial_state, PC_JACOBI, PC_IC0,
                          PC_LEGACY_IC_A, PC_CHEBYSHEV, PC_NONE)

pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build absent")


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.set_threads(1)
    return o


@pytest.fixture(scope="module")
def ref():
    r = Reference()
    r.set_threads(1)
    return r


def case(nx, ny, nz, seed, frac=0.08):
    m = dict(nx=nx, ny=ny, nz=nz, dx=1.0 / nx, dy=0.7 / ny, dz=1.3 / nz)
    N = nx * ny * nz
    solid = (wl.uniform(seed, 9, N, 0, 1) < frac).astype(np.uint8)
    f = [wl.uniform(seed, s, N) for s in range(8)]
    return m, solid, f


SHAPES = [(7, 6, 5, 1), (9, 8, 11, 2), (12, 10, 9, 3)]


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("shape", SHAPES)
def test_stencil_and_matvec(orc, ref, shape):
    m, solid, f = case(*shape)
    A_o, A_r = orc.build_stencil(solid, m), ref.build_stencil(solid, m)
    for a, b in zip(A_o, A_r):
        assert same(a, b)
    assert same(orc.matvec(f[0], A_r, m), ref.matvec(f[0], A_r, m))


@pytest.mark.parametrize("shape", SHAPES)
def test_divergence_b(orc, ref, shape):
    m, solid, f = case(*shape)
    d_o = orc.divergence(f[0], f[1], f[2], solid, m, 1e-3, 1.2)
    d_r = ref.divergence(f[0], f[1], f[2], solid, m, 1e-3, 1.2)
    assert same(d_o, d_r)
    assert same(orc.build_b(d_o, solid, m), ref.build_b(d_r, solid, m))


@pytest.mark.parametrize("shape", SHAPES)
def test_preconditioners(orc, ref, shape):
    m, solid, f = case(*shape)
    A = ref.build_stencil(solid, m)
    assert same(orc.jacobi_precond(f[0], A[0]), ref.jacobi_precond(f[0], A[0]))
    L_o, L_r = orc.ic_build(A, m), ref.ic_build(A, m)
    for a, b in zip(L_o, L_r):
        assert same(a, b)
    assert same(orc.ic_apply(f[1], L_r, m), ref.ic_apply(f[1], L_r, m))
    assert same(orc.ic_apply(f[1], A, m), ref.ic_apply(f[1], A, m))
    lmax = ref.lambda_max(A, m, 20)
    assert orc.lambda_max(A, m, 20) == lmax
    for deg in (1, 2, 5, 8):
        assert same(orc.cheb_apply(f[2], A, m, 3 * np.pi ** 2, lmax, deg),
                    ref.cheb_apply(f[2], A, m, 3 * np.pi ** 2, lmax, deg))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("pc", [PC_NONE, PC_JACOBI, PC_IC0, PC_LEGACY_IC_A, PC_CHEBYSHEV])
def test_pcg_trajectory(orc, ref, shape, pc):
    m, solid, f = case(*shape)
    A = ref.build_stencil(solid, m)
    L = ref.ic_build(A, m)
    lmax = ref.lambda_max(A, m, 20)
    b = ref.build_b(f[0], solid, m)
    tol = 1e-10 * np.linalg.norm(b)
    xo, io, to = orc.pcg(A, b, f[1], m, pc, L=L, lmin=3 * np.pi ** 2, lmax=lmax, degree=4,
                         max_iter=400, tol=tol)
    xr, ir, tr = ref.pcg(A, b, f[1], m, pc, L=L, lmin=3 * np.pi ** 2, lmax=lmax, degree=4,
                         max_iter=400, tol=tol)
    assert io == ir
    assert same(to, tr)
    assert same(xo, xr)


def test_harness_loop_equals_shipped_pcg(ref):
    """The harness PCG loop with IC-applied-with-A reproduces
    solve_pressure_pcg bit for bit, and with Chebyshev reproduces
    solve_pressure_pcg_chebyshev."""
    m, solid, f = case(9, 8, 11, 5)
    A = ref.build_stencil(solid, m)
    b = ref.build_b(f[0], solid, m)
    x1, i1 = ref.pcg_legacy(A, b, f[1], m, max_iter=300, tol=1e-9)
    x2, i2, _ = ref.pcg(A, b, f[1], m, PC_LEGACY_IC_A, max_iter=300, tol=1e-9)
    assert i1 == i2 and same(x1, x2)
    lmax = ref.lambda_max(A, m, 20)
    x1, i1 = ref.pcg_cheb(A, b, f[1], m, 3 * np.pi ** 2, lmax, 5, 300, 1e-9)
    x2, i2, _ = ref.pcg(A, b, f[1], m, PC_CHEBYSHEV, lmin=3 * np.pi ** 2, lmax=lmax, degree=5,
                        max_iter=300, tol=1e-9)
    assert i1 == i2 and same(x1, x2)


@pytest.mark.parametrize("shape", SHAPES)
def test_stationary(orc, ref, shape):
    m, solid, f = case(*shape)
    p0 = f[3] * 0.01
    assert same(orc.pressure_bcs(f[3], solid, m), ref.pressure_bcs(f[3], solid, m))
    po, _ = orc.jacobi_solve(p0, f[0], f[1], f[2], solid, m, 1e-3, 1.2)
    pr, _ = ref.jacobi_solve(p0, f[0], f[1], f[2], solid, m, 1e-3, 1.2)
    assert same(po, pr)
    for it in (1, 7, 200):
        po, _ = orc.gs_solve(p0, f[0], f[1], f[2], solid, m, 1e-3, 1.2, it)
        pr, _ = ref.gs_solve(p0, f[0], f[1], f[2], solid, m, 1e-3, 1.2, it)
        assert same(po, pr)


@pytest.mark.parametrize("shape", SHAPES)
def test_physics(orc, ref, shape):
    m, solid, f = case(*shape)
    u, v, w = f[0], f[1], f[2]
    nt0 = np.abs(f[3]) * 1e-3
    assert orc.adaptive_dt(u, v, w, nt0, m, 1e-3) == ref.adaptive_dt(u, v, w, nt0, m, 1e-3)
    no, to = orc.sgs(u, v, w, m, 0.18, nu_t=nt0, want_tau=True)
    nr, tr = ref.sgs(u, v, w, m, 0.18, nu_t=nt0, want_tau=True)
    assert same(no, nr)
    for a, b in zip(to, tr):
        assert same(a, b)
    assert same(orc.nut_bcs(nr, solid, m), ref.nut_bcs(nr, solid, m))
    mo = orc.momentum(u, v, w, f[4], f[5], f[6], nr, m, 1e-3, 1e-3)
    mr = ref.momentum(u, v, w, f[4], f[5], f[6], nr, m, 1e-3, 1e-3)
    for a, b in zip(mo, mr):
        assert same(a, b)
    for a, b in zip(orc.velocity_bcs(u, v, w, solid, m), ref.velocity_bcs(u, v, w, solid, m)):
        assert same(a, b)
    for a, b in zip(orc.correct(u, v, w, f[7], m, 1e-3, 1.2),
                    ref.correct(u, v, w, f[7], m, 1e-3, 1.2)):
        assert same(a, b)


# oracle precond code -> reference harness solver code
SOLVER = {PC_LEGACY_IC_A: 0, PC_JACOBI: 1, PC_IC0: 2, PC_CHEBYSHEV: 4}


@pytest.mark.parametrize("pc", [PC_LEGACY_IC_A, PC_JACOBI, PC_IC0, PC_CHEBYSHEV])
def test_steps(orc, ref, pc):
    """Whole steps (main.cpp:115-144), including the lagged projection,
    ping-pong faces and inlet solids, bit for bit."""
    m, solid, f = case(10, 9, 8, 11, frac=0.06)
    solid = solid.copy()
    solid[[3, 17, 40]] = 1  # inlet-face solid cells (i = 0)
    N = solid.size
    st0 = initial_state(f[0], f[1], f[2], N, solid, m, ref)
    A = ref.build_stencil(solid, m)
    L = ref.ic_build(A, m) if pc == PC_IC0 else None
    lmax = ref.lambda_max(A, m, 20)
    kw = dict(nu=1e-3, rho=1.2, lmin=3 * np.pi ** 2, lmax=lmax, degree=4, max_iter=500,
              tol=1e-8, tol_rel=True)
    so, io, do = orc.run_steps(st0, solid, m, A, 4, step0=19, precond=pc, L=L, **kw)
    sr, ir, dr = ref.run_steps(st0, solid, m, A, 4, step0=19, solver=SOLVER[pc], **kw)
    assert io == ir and do == dr
    for k in so:
        assert same(so[k], sr[k]), k


def test_steps_gs(orc, ref):
    """The SOR alternative (main.cpp:133) over steps."""
    m, solid, f = case(9, 8, 7, 4, frac=0.05)
    st0 = initial_state(f[0], f[1], f[2], solid.size, solid, m, ref)
    sr, _, _ = ref.run_steps(st0, solid, m, None, 2, solver=5, gs_iters=30)
    # the oracle composes the same sequence by hand
    st = {k: st0[k].copy() for k in st0}
    for _ in range(2):
        dt, rho = 1e-3, 1.2
        st["nu_t"] = orc.nut_bcs(orc.sgs(st["u"], st["v"], st["w"], m, 0.18, nu_t=st["nu_t"]),
                                 solid, m)
        st["u"], st["v"], st["w"], st["u_new"], st["v_new"], st["w_new"] = orc.momentum(
            st["u"], st["v"], st["w"], st["u_new"], st["v_new"], st["w_new"], st["nu_t"], m, dt,
            1e-3)
        st["u"], st["v"], st["w"] = orc.velocity_bcs(st["u"], st["v"], st["w"], solid, m)
        st["p"], _ = orc.gs_solve(st["p"], st["u"], st["v"], st["w"], solid, m, dt, rho, 30)
        st["u"], st["v"], st["w"] = orc.correct(st["u"], st["v"], st["w"], st["p"], m, dt, rho)
        st["u"], st["v"], st["w"] = orc.velocity_bcs(st["u"], st["v"], st["w"], solid, m)
    for k in ("u", "v", "w", "u_new", "v_new", "w_new", "p", "nu_t"):
        assert same(st[k], sr[k]), k
