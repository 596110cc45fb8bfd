"""The C++ drop-in layer (include/les/pressure.h, include/les/physics.h,
les-model_amd/liblespressure_gpu.so): a C++ program written against the
reference's function signatures (tests/cpp/dropin_check.cpp) runs on the GPU
and its outputs are checked against the oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from cases import bits_equal, rel_err
from cpu_checkers import Oracle, PC_LEGACY_IC_A, PC_CHEBYSHEV

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "les-model_amd", "build", "dropin_check")


def gen(n, seed):
    i = np.arange(n, dtype=np.int64)
    return ((i * 7919 + seed * 104729) % 2001) / 1000.0 - 1.0


def read_dump(path):
    out = {}
    with open(path, "rb") as fh:
        while True:
            nm = fh.read(32)
            if not nm:
                break
            n = int(np.frombuffer(fh.read(8), dtype=np.int64)[0])
            out[nm.rstrip(b"\0").decode()] = np.frombuffer(fh.read(8 * n), dtype=np.float64)
    return out


def test_dropin_program(tmp_path):
    assert os.path.exists(EXE), "build the library first (make -C les-model_amd)"
    dump = tmp_path / "dropin.bin"
    subprocess.run([EXE, str(dump)], check=True, timeout=120)
    D = read_dump(dump)
    m = dict(nx=12, ny=10, nz=9, dx=1.0 / 12, dy=0.7 / 10, dz=1.3 / 9)
    N = 12 * 10 * 9
    u, v, w, un, vn, wn, p = (gen(N, s) for s in range(1, 8))
    nt = (gen(N, 8) + 1.0) * 1e-3
    solid = (((np.arange(N) * 31) % 17) == 0).astype(np.uint8)
    o = Oracle()
    o.set_threads(1)
    A = o.build_stencil(solid, m)
    for q, name in enumerate(("A_c", "A_n", "A_s", "A_e", "A_w", "A_t", "A_b")):
        assert bits_equal(D[name], A[q]), name
    assert bits_equal(D["matvec"], o.matvec(u, A, m))
    div = o.divergence(un, vn, wn, solid, m, 1e-3, 1.2)
    assert bits_equal(D["divergence"], div)
    b = o.build_b(div, solid, m)
    assert bits_equal(D["b"], b)
    assert bits_equal(D["jacobi"], o.jacobi_precond(v, A[0]))
    L = o.ic_build(A, m)
    for q, name in ((0, "L_c"), (4, "L_w"), (2, "L_s"), (6, "L_b")):
        assert bits_equal(D[name], L[q]), name
    assert bits_equal(D["ic_L"], o.ic_apply(v, L, m))
    lmax = o.lambda_max(A, m, 20)
    assert abs(D["lambda_max"][0] - lmax) <= 1e-12 * lmax
    lmax_g = D["lambda_max"][0]
    assert bits_equal(D["cheb5"], o.cheb_apply(v, A, m, 3 * np.pi ** 2, lmax_g, 5))
    x, it, _ = o.pcg(A, b, w, m, PC_LEGACY_IC_A, max_iter=20, tol=0.0)
    assert D["pcg_legacy_20_iters"][0] == it == 20
    assert rel_err(D["pcg_legacy_20"], x) < 1e-8
    x, it, _ = o.pcg(A, b, w, m, PC_CHEBYSHEV, lmin=3 * np.pi ** 2, lmax=lmax_g, degree=5,
                     max_iter=20, tol=0.0)
    assert D["pcg_cheb_20_iters"][0] == it
    assert rel_err(D["pcg_cheb_20"], x) < 1e-10
    assert D["dot"][0] == o.dot(u, v)
    assert bits_equal(D["pressure_bcs"], o.pressure_bcs(p, solid, m))
    assert bits_equal(D["jacobi_solve"], o.jacobi_solve(p * 0.01, u, v, w, solid, m, 1e-3, 1.2)[0])
    assert bits_equal(D["gs_10"], o.gs_solve(p * 0.01, u, v, w, solid, m, 1e-3, 1.2, 10)[0])
    assert D["adaptive_dt"][0] == o.adaptive_dt(u, v, w, nt, m, 1e-3)
    nto, tau = o.sgs(u, v, w, m, 0.18, nu_t=nt, want_tau=True)
    assert bits_equal(D["sgs_nu_t"], nto)
    for q, name in enumerate(("tau_xx", "tau_xy", "tau_xz", "tau_yy", "tau_yz", "tau_zz")):
        assert bits_equal(D[name], tau[q]), name
    ntb = o.nut_bcs(nto, solid, m)
    assert bits_equal(D["nut_bcs"], ntb)
    mo = o.momentum(u, v, w, un, vn, wn, ntb, m, 1e-3, 1e-3)
    for q, name in enumerate(("mom_u", "mom_v", "mom_w", "mom_u_new", "mom_v_new", "mom_w_new")):
        assert bits_equal(D[name], mo[q]), name
    vb = o.velocity_bcs(mo[0], mo[1], mo[2], solid, m)
    for q, name in enumerate(("vbc_u", "vbc_v", "vbc_w")):
        assert bits_equal(D[name], vb[q]), name
    co = o.correct(vb[0], vb[1], vb[2], p, m, 1e-3, 1.2)
    for q, name in enumerate(("cor_u", "cor_v", "cor_w")):
        assert bits_equal(D[name], co[q]), name
