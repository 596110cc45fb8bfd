"""The multigrid preconditioner on the GPU (no reference counterpart: the
reference's README.md:53 lists multigrid as a TODO) against the oracle's
restatement (oracle/les_oracle.c, pinned by tests/test_oracle_mg.py):
the FP64 V-cycle bit for bit, PCG with it to the oracle's iterates (in the
Euclidean and in the mirror-weighted inner product), the coded operator
identical to the arrays, mixed precision to the FP64 solution, and whole
LES steps."""
import numpy as np
import pytest

import lesgpu
import les_workloads as wl
from cases import random_case, box_case, bits_equal, rel_err
from cpu_checkers import Oracle, initial_state, PC_MG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.set_threads(1)
    return o


def ctx_for(m, solid=None):
    g = lesgpu.LesGpu.from_mesh(m)
    if solid is not None:
        g.set_mask(solid)
    return g


SHAPES = [(7, 6, 5, 1), (13, 9, 17, 2), (4, 3, 1032, 5), (9, 13, 64, 7), (64, 48, 40, 8),
          (3, 4, 3, 9)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", ["mask", "free", "general"])
def test_vcycle_bitexact(orc, shape, kind):
    """One V-cycle (les_precond_apply, FP64, arrays) equals the oracle's bit
    for bit: the coarse operators, sweeps, restriction and prolongation
    follow its operation order.  'general' is a StencilMatrix with every
    coefficient set (through les_set_stencil)."""
    m, solid, f = random_case(*shape, frac=0.1)
    print(f"[mg] V-cycle {shape} {kind}", flush=True)
    with ctx_for(m, solid) as g:
        if kind == "general":
            A = [f[q] + (4.0 if q == 0 else 0.0) for q in range(7)]
            g.set_stencil(A)
        else:
            A = orc.build_stencil(solid if kind == "mask" else np.zeros_like(solid), m)
            g.build_stencil(ignore_mask=(kind == "free"))
        for sweeps, coarse in ((2, 32), (1, 7)):
            g.set_option(lesgpu.OPT_MG_SWEEPS, sweeps)
            g.set_option(lesgpu.OPT_MG_COARSE_SWEEPS, coarse)
            z = g.precond_apply(lesgpu.PC_MG, f[3])
            ref = orc.mg_apply(f[3], A, m, sweeps=sweeps, coarse_sweeps=coarse)
            assert bits_equal(z, ref), (sweeps, coarse, rel_err(z, ref))


def test_hierarchy_follows_the_operator(orc):
    """The coarse levels are rebuilt when A changes (les_build_stencil,
    les_set_stencil)."""
    m, solid, f = random_case(16, 12, 20, 3, frac=0.1)
    A1 = orc.build_stencil(solid, m)
    A2 = [a * 1.5 for a in A1]
    with ctx_for(m, solid) as g:
        g.build_stencil()
        assert bits_equal(g.precond_apply(lesgpu.PC_MG, f[0]), orc.mg_apply(f[0], A1, m))
        g.set_stencil(A2)
        assert bits_equal(g.precond_apply(lesgpu.PC_MG, f[0]), orc.mg_apply(f[0], A2, m))
        g.build_stencil()
        assert bits_equal(g.precond_apply(lesgpu.PC_MG, f[0]), orc.mg_apply(f[0], A1, m))


@pytest.mark.parametrize("shape", [(20, 16, 24, 4), (33, 20, 70, 3), (64, 48, 40, 8)])
def test_pcg_mg_fp64(orc, shape):
    """PCG-MG in FP64: fixed-iteration solves follow the oracle's iterates to
    1e-10 (only the dot products' summation order differs), solves to
    1e-8 relative take the oracle's iteration count +-1, and the coded
    operator gives the arrays' solve bit for bit."""
    m, solid, f = random_case(*shape, frac=0.05)
    A = orc.build_stencil(solid, m)
    b = orc.build_b(orc.divergence(f[0], f[1], f[2], solid, m, 1e-3, 1.2), solid, m)
    x0 = f[4] * 0.1
    xr, _, tr = orc.pcg_mg(A, b, x0, m, max_iter=6, tol=0.0)
    tol = 1e-8 * np.linalg.norm(b)
    xt, itr, _ = orc.pcg_mg(A, b, np.zeros_like(b), m, max_iter=500, tol=tol)
    print(f"[mg] PCG {shape}: oracle {itr} iterations", flush=True)
    outs = []
    with ctx_for(m, solid) as g:
        g.build_stencil()
        g.set_field("b", b)
        for fmt in (lesgpu.STENCIL_ARRAYS, lesgpu.STENCIL_CODED):
            g.set_field("p", x0)
            r = g.pcg(precond=lesgpu.PC_MG, max_iter=6, tol=0.0, stencil_format=fmt)
            x = g.get_field("p")
            assert rel_err(x, xr) < 1e-10, rel_err(x, xr)
            assert np.allclose(g.pcg_trace(7), tr, rtol=1e-9)
            g.set_field("p", np.zeros_like(b))
            r2 = g.pcg(precond=lesgpu.PC_MG, max_iter=500, tol=1e-8, tol_rel=True,
                       stencil_format=fmt)
            assert r2["converged"] == 1 and abs(r2["iters"] - itr) <= 1, (r2, itr)
            x2 = g.get_field("p")
            assert np.linalg.norm(b - orc.matvec(x2, A, m)) < 1.01 * tol
            assert rel_err(x2, xt) < 1e-7
            outs.append((r, x, r2, x2))
    (ra, xa, r2a, x2a), (rb, xb, r2b, x2b) = outs
    assert ra == rb and r2a == r2b and bits_equal(xa, xb) and bits_equal(x2a, x2b)


@pytest.mark.parametrize("shape,masked", [((48, 40, 64, 6), True), ((64, 64, 64, 0), False)])
def test_pcg_mg_mixed(orc, shape, masked):
    """Mixed-precision PCG-MG (FP32 V-cycle and Krylov vectors, FP64
    reductions and residual replacement) meets 1e-8 on the FP64 residual,
    agrees with the FP64 oracle solution to 1e-5, and its coded and arrays
    solves are identical."""
    m, solid, f = random_case(*shape, frac=0.05)
    A = orc.build_stencil(solid if masked else np.zeros_like(solid), m)
    b = orc.build_b(orc.divergence(f[0], f[1], f[2], solid, m, 1e-3, 1.2), solid, m)
    x_ref, it_ref, _ = orc.pcg_mg(A, b, np.zeros_like(b), m, max_iter=500,
                                  tol=1e-13 * np.linalg.norm(b))
    outs = []
    with ctx_for(m, solid) as g:
        g.build_stencil(ignore_mask=not masked)
        g.set_field("b", b)
        for fmt in (lesgpu.STENCIL_ARRAYS, lesgpu.STENCIL_CODED):
            g.set_field("p", np.zeros_like(b))
            r = g.pcg(precond=lesgpu.PC_MG, max_iter=500, tol=1e-8, tol_rel=True,
                      precision=lesgpu.PREC_MIXED, reliable_every=50, stencil_format=fmt)
            x = g.get_field("p")
            print(f"[mg] mixed {shape} fmt {fmt}: {r['iters']} iterations (FP64 oracle to 1e-13: "
                  f"{it_ref})", flush=True)
            assert r["converged"] == 1 and r["fault"] == 0, r
            assert np.linalg.norm(b - orc.matvec(x, A, m)) < 1.01e-8 * np.linalg.norm(b)
            assert rel_err(x, x_ref) < 1e-5
            outs.append((r, x))
    assert outs[0][0] == outs[1][0] and bits_equal(outs[0][1], outs[1][1])


def test_steps_mg(orc):
    """Whole LES steps with PCG-MG in the mirror-weighted inner product
    (main.cpp:108-165 with the preconditioner swapped; A built before the
    mask, as main.cpp:65-81 builds it) follow the oracle's steps, from a
    U(-1,1) velocity field (the first step's right-hand side is white noise,
    where the Euclidean iteration with the unweighted cycle stalls)."""
    n = 24
    m, solid, f = box_case(n, 5)
    A = orc.build_stencil(np.zeros_like(solid), m)
    W = orc.mirror_weights(A, m)
    st0 = initial_state(f[0], f[1], f[2], n ** 3, solid, m, orc)
    kw = dict(max_iter=500, tol=1e-10, tol_rel=True)
    ref, iters, _ = orc.run_steps(st0, solid, m, A, 3, precond=PC_MG, nu=1e-3, rho=1.2, Cs=0.17,
                                  W=W, **kw)
    assert max(iters) < 100, iters
    for prec in (lesgpu.PREC_FP64, lesgpu.PREC_MIXED):
        with ctx_for(m, solid) as g:
            g.set_state(st0)
            g.build_stencil(ignore_mask=True)
            gi = []
            for s in range(3):
                r = g.step(step=s, nu=1e-3, rho=1.2, Cs=0.17, precond=lesgpu.PC_MG,
                           precision=prec, inner_product=lesgpu.IP_MIRROR, **kw)
                assert r["converged"] == 1, r
                gi.append(r["iters"])
            out = g.get_state()
        print(f"[mg] steps precision {prec}: iterations gpu {gi} oracle {iters}", flush=True)
        if prec == lesgpu.PREC_FP64:
            assert all(abs(a - b) <= 1 for a, b in zip(gi, iters))
            for k in ("u", "v", "w", "p", "nu_t"):
                assert rel_err(out[k], ref[k]) < 1e-8, (k, rel_err(out[k], ref[k]))
        else:
            for k in ("u", "v", "w", "p"):
                assert rel_err(out[k], ref[k]) < 1e-6, (k, rel_err(out[k], ref[k]))


@pytest.mark.parametrize("shape", [(20, 16, 24, 4), (64, 48, 40, 8)])
def test_pcg_mg_weighted(orc, shape):
    """PCG-MG in the mirror-weighted inner product on the operator the
    shipped driver builds (no interior obstacles): FP64 iterates follow the
    oracle's to 1e-10 at a fixed iteration count and its iteration count
    +-1 to 1e-8; mixed precision meets 1e-8 in FP64 and the FP64 solution to
    1e-5; coded and arrays identical."""
    m, solid, f = random_case(*shape, frac=0.05)
    A = orc.build_stencil(np.zeros_like(solid), m)
    W = orc.mirror_weights(A, m)
    b = orc.build_b(orc.divergence(f[0], f[1], f[2], solid, m, 1e-3, 1.2), solid, m)
    x0 = f[4] * 0.1
    xr, _, tr = orc.pcg_mg(A, b, x0, m, max_iter=6, tol=0.0, W=W)
    tol = 1e-8 * np.linalg.norm(b)
    xt, itr, _ = orc.pcg_mg(A, b, np.zeros_like(b), m, max_iter=500, tol=tol, W=W)
    print(f"[mg] weighted PCG {shape}: oracle {itr} iterations", flush=True)
    outs = []
    with ctx_for(m, solid) as g:
        g.build_stencil(ignore_mask=True)
        g.set_field("b", b)
        for fmt in (lesgpu.STENCIL_ARRAYS, lesgpu.STENCIL_CODED):
            kw = dict(precond=lesgpu.PC_MG, stencil_format=fmt, inner_product=lesgpu.IP_MIRROR)
            g.set_field("p", x0)
            r = g.pcg(max_iter=6, tol=0.0, **kw)
            x = g.get_field("p")
            assert rel_err(x, xr) < 1e-10, rel_err(x, xr)
            assert np.allclose(g.pcg_trace(7), tr, rtol=1e-9)
            g.set_field("p", np.zeros_like(b))
            r2 = g.pcg(max_iter=500, tol=1e-8, tol_rel=True, **kw)
            assert r2["converged"] == 1 and abs(r2["iters"] - itr) <= 1, (r2, itr)
            x2 = g.get_field("p")
            assert rel_err(x2, xt) < 1e-7
            g.set_field("p", np.zeros_like(b))
            r3 = g.pcg(max_iter=500, tol=1e-8, tol_rel=True, precision=lesgpu.PREC_MIXED,
                       reliable_every=50, **kw)
            x3 = g.get_field("p")
            assert r3["converged"] == 1 and r3["fault"] == 0, r3
            assert np.linalg.norm(b - orc.matvec(x3, A, m)) < 1.01e-8 * np.linalg.norm(b)
            assert rel_err(x3, xt) < 1e-5
            # the residual is replaced after every 3 decades of drop (before
            # the FP32 drift matters), so the direction never restarts and
            # mixed precision takes a few iterations more than FP64 at most
            # (a replacement keeps the direction, not all of its conjugacy;
            # measured: 3 more at 64x48x40, 0.9 more a step at 256^3)
            assert r3["restarts"] == 0 and r3["replacements"] >= 2, r3
            assert r3["iters"] <= r2["iters"] + 5, (r3, r2)
            outs.append((r, x, r2, x2, r3, x3))
    a, b_ = outs
    assert a[0] == b_[0] and a[2] == b_[2] and a[4] == b_[4]
    assert bits_equal(a[1], b_[1]) and bits_equal(a[3], b_[3]) and bits_equal(a[5], b_[5])


def test_mg_options_range():
    m, solid, _ = random_case(4, 8, 8, 1, frac=0.0)
    with ctx_for(m, solid) as g:
        for opt, bad in ((lesgpu.OPT_MG_SWEEPS, (0, 17)), (lesgpu.OPT_MG_COARSE_SWEEPS, (0, 4097))):
            for v in bad:
                with pytest.raises(lesgpu.LesError, match="multigrid"):
                    g.set_option(opt, v)
        for v in (-1, 16):
            with pytest.raises(lesgpu.LesError, match="replacement decades"):
                g.set_option(lesgpu.OPT_REPLACE_DECADES, v)
        g.build_stencil()
        with pytest.raises(lesgpu.LesError, match="multigrid"):
            g.pcg(precond=lesgpu.PC_CHEBYSHEV, max_iter=3, inner_product=lesgpu.IP_MIRROR,
                  precision=lesgpu.PREC_MIXED)


@pytest.mark.parametrize("precision", [lesgpu.PREC_FP64, lesgpu.PREC_MIXED])
@pytest.mark.parametrize("precond", [PC_MG, lesgpu.PC_CHEBYSHEV])
def test_host_batches_change_nothing(orc, precision, precond):
    """How many iterations the host issues between its polls (fixed batches
    of 1 or 3, or the default batches that stop at the predicted last
    iteration) changes neither the iterates nor the counts: the device
    decides when the solve is done.  Mixed precision with Chebyshev is the
    exception in its last bits: the host enqueues the gated FP64
    replacement only periodically until its polls see the residual near the
    tolerance, so a pending replacement runs at an iteration that depends on
    the poll schedule (reproducibly for a given schedule); the counts stay
    the same and the results agree to well inside the tolerance.  A
    multigrid solve enqueues the replacement on every iteration."""
    m, solid, f = random_case(40, 36, 64, 5, frac=0.05)
    A = orc.build_stencil(np.zeros_like(solid), m)
    b = orc.build_b(orc.divergence(f[0], f[1], f[2], solid, m, 1e-3, 1.2), solid, m)
    lmax = orc.lambda_max(A, m, 20)
    kw = dict(precond=precond, max_iter=2000, tol=1e-9, tol_rel=True, precision=precision,
              reliable_every=50, lmin=3 * np.pi ** 2, lmax=1.05 * lmax, degree=8,
              stencil_format=lesgpu.STENCIL_CODED,
              inner_product=lesgpu.IP_MIRROR if precond == PC_MG else lesgpu.IP_EUCLID)
    outs = []
    with ctx_for(m, solid) as g:
        g.build_stencil(ignore_mask=True)
        g.set_field("b", b)
        for check in (0, 1, 3):
            g.set_field("p", np.zeros_like(b))
            r = g.pcg(check_every=check, **kw)
            assert r["converged"] == 1 and r["fault"] == 0, (check, r)
            outs.append((r, g.get_field("p")))
    print(f"[mg] batches precision {precision} precond {precond}: {outs[0][0]['iters']} iterations",
          flush=True)
    exact = precision == lesgpu.PREC_FP64 or precond == PC_MG
    r0, x0 = outs[0]
    for r, x in outs[1:]:
        if exact:
            assert r == r0
            assert bits_equal(x, x0)
        else:
            assert (r["iters"], r["replacements"], r["restarts"]) == \
                (r0["iters"], r0["replacements"], r0["restarts"])
            assert r["true_res_norm"] < 1e-9 * r["b_norm"]
            assert rel_err(x, x0) < 1e-5


@pytest.mark.parametrize("shape", [(20, 24, 64, 21), (9, 40, 128, 22), (12, 30, 256, 23),
                                   (17, 13, 16, 24), (7, 9, 36, 25), (4, 6, 2048, 26)])
@pytest.mark.parametrize("ip", [lesgpu.IP_EUCLID, lesgpu.IP_MIRROR])
def test_vcycle_fp32_rings_identical(orc, shape, ip):
    """The FP32 V-cycle, as a mixed-precision PCG applies it: the one-pass
    pre- and post-smoothing rings give bit for bit the separate sweeps'
    result, with the code and with the arrays, in the Euclidean and the
    mirror-weighted cycle; and it follows the FP64 V-cycle to FP32 rounding.
    nz = 64 .. 256 take the rings at level 0 with a few rows per workgroup,
    nz = 16 with hundreds, nz = 36 (not a multiple of 8) the scalar kernels
    below level 0; nz = 2048 has rows too long for the two-ring passes, and
    its level 1 (nz = 1024) takes the restriction ring with two core rows."""
    m, solid, f = random_case(*shape, frac=0.05)
    r = f[3].astype(np.float32)
    print(f"[mg] FP32 V-cycle {shape} ip {ip}", flush=True)
    with ctx_for(m, solid) as g:
        g.build_stencil(ignore_mask=True)
        outs = []
        for rings in (1, 0):
            g.set_option(lesgpu.OPT_MG_RINGS, rings)
            for fmt in (lesgpu.STENCIL_CODED, lesgpu.STENCIL_ARRAYS):
                outs.append(g.precond_apply_fp32(PC_MG, r, stencil_format=fmt, inner_product=ip))
        for z in outs[1:]:
            assert bits_equal(z, outs[0])
        assert np.all(np.isfinite(outs[0]))
        if ip == lesgpu.IP_EUCLID:
            z64 = g.precond_apply(PC_MG, r.astype(np.float64))
            assert rel_err(outs[0].astype(np.float64), z64) < 1e-4
