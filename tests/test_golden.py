"""The oracle against the committed golden fixtures (tests/golden/), written by
the unmodified reference solver (tests/golden/make_golden.py).  Runs on the
CPU everywhere, including where the reference build is absent; bit for bit
on one thread."""
import json
import os

import numpy as np
import pytest

import les_workloads as wl
from cases import bits_equal
from cpu_checkers import Oracle, STATE_KEYS, initial_state, PC_JACOBI, PC_LEGACY_IC_A

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.set_threads(1)
    return o


@pytest.fixture(scope="module")
def K():
    return dict(np.load(os.path.join(GOLD, "kernels_10x9x8.npz")))


def test_kernels(orc, K):
    m = dict(nx=int(K["nx"]), ny=int(K["ny"]), nz=int(K["nz"]), dx=float(K["dx"]),
             dy=float(K["dy"]), dz=float(K["dz"]))
    solid = K["solid"]
    f = [K[f"in{q}"] for q in range(8)]
    A = orc.build_stencil(solid, m)
    assert bits_equal(np.stack(A), K["A"])
    assert bits_equal(orc.matvec(f[0], A, m), K["matvec"])
    div = orc.divergence(f[0], f[1], f[2], solid, m, 1e-3, 1.2)
    assert bits_equal(div, K["divergence"])
    assert bits_equal(orc.build_b(div, solid, m), K["b"])
    assert bits_equal(orc.jacobi_precond(f[3], A[0]), K["jacobi"])
    L = orc.ic_build(A, m)
    assert bits_equal(np.stack(L), K["L"])
    assert bits_equal(orc.ic_apply(f[3], L, m), K["ic_apply_L"])
    assert bits_equal(orc.ic_apply(f[3], A, m), K["ic_apply_A"])
    lmax = orc.lambda_max(A, m, 20)
    assert lmax == K["lambda_max"][0]
    assert bits_equal(orc.cheb_apply(f[3], A, m, 3 * np.pi ** 2, lmax, 5), K["cheb5"])
    x, it, _ = orc.pcg(A, K["b"], f[4], m, PC_LEGACY_IC_A, max_iter=30, tol=0.0)
    assert bits_equal(x, K["pcg_legacy_30"])
    x, it, _ = orc.pcg(A, K["b"], f[4], m, 4, lmin=3 * np.pi ** 2, lmax=lmax, degree=5,
                       max_iter=30, tol=0.0)
    assert bits_equal(x, K["pcg_cheb5_30"])
    assert bits_equal(orc.pressure_bcs(f[5], solid, m), K["pressure_bcs"])
    p0 = f[5] * 0.01
    assert bits_equal(orc.jacobi_solve(p0, f[0], f[1], f[2], solid, m, 1e-3, 1.2)[0],
                      K["jacobi_solve"])
    assert bits_equal(orc.gs_solve(p0, f[0], f[1], f[2], solid, m, 1e-3, 1.2, 10)[0],
                      K["gs_solve_10"])
    nt0 = np.abs(f[6]) * 1e-3
    assert orc.adaptive_dt(f[0], f[1], f[2], nt0, m, 1e-3) == K["adaptive_dt"][0]
    nt, tau = orc.sgs(f[0], f[1], f[2], m, 0.18, nu_t=nt0, want_tau=True)
    assert bits_equal(nt, K["sgs_nu_t"])
    assert bits_equal(np.stack(tau), K["sgs_tau"])
    ntb = orc.nut_bcs(nt, solid, m)
    assert bits_equal(ntb, K["nut_bcs"])
    mo = orc.momentum(f[0], f[1], f[2], f[5], f[6], f[7], ntb, m, 1e-3, 1e-3)
    assert bits_equal(np.stack(mo), K["momentum"])
    assert bits_equal(np.stack(orc.velocity_bcs(mo[0], mo[1], mo[2], solid, m)),
                      K["velocity_bcs"])
    assert bits_equal(np.stack(orc.correct(f[0], f[1], f[2], f[4], m, 1e-3, 1.2)), K["correct"])


def test_steps16(orc):
    S = dict(np.load(os.path.join(GOLD, "steps_16.npz")))
    n = int(S["n"])
    m = wl.mesh(n)
    A = orc.build_stencil(S["solid"], m)
    st0 = {k: S["in_" + k] for k in STATE_KEYS}
    out, its, dts = orc.run_steps(st0, S["solid"], m, A, 3, step0=int(S["step0"]),
                                  precond=PC_LEGACY_IC_A, nu=1e-3, rho=1.2, Cs=0.18,
                                  max_iter=1000, tol=1e-6)
    assert list(its) == list(S["iters"])
    assert list(dts) == list(S["dts"])
    for k in STATE_KEYS:
        assert bits_equal(out[k], S["out_" + k]), k


def test_c1_digest(orc):
    """The CPU-runnable configuration: 32^3, 10 steps, PCG-Jacobi to 1e-8,
    A built before the mask (main.cpp:65-81)."""
    with open(os.path.join(GOLD, "c1_digest.json")) as fh:
        D = json.load(fh)
    m, (u, v, w), solid = wl.config_c1()
    A = orc.build_stencil(np.zeros_like(solid), m)
    st0 = initial_state(u, v, w, 32 ** 3, solid, m, orc)
    out, its, dts = orc.run_steps(st0, solid, m, A, 10, precond=PC_JACOBI, nu=1e-3, rho=1.2,
                                  Cs=0.18, max_iter=20000, tol=1e-8, tol_rel=True)
    assert list(its) == D["iters"]
    assert list(dts) == D["dts"]
    import hashlib
    for k in STATE_KEYS:
        a = np.ascontiguousarray(out[k])
        assert hashlib.sha256(a.tobytes()).hexdigest() == D["fields"][k]["sha256"], k
