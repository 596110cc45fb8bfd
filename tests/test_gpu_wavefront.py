"""The lexicographic sweeps as single wavefront launches (IC(0) factor and
solves, lexicographic SOR): bit for bit against one launch per hyperplane
and against the oracle, repeated launches, and the host-set spin limit."""
import numpy as np
import pytest

import lesgpu
from cases import random_case, box_case, bits_equal
from cpu_checkers import Oracle

pytestmark = pytest.mark.gpu

# ragged tiles in i and j (8 x 8 columns per tile), one tile, many tiles
SHAPES = [(3, 3, 3, 1), (8, 8, 5, 2), (9, 17, 6, 3), (33, 20, 70, 4), (41, 40, 9, 5),
          (5, 70, 33, 6), (64, 64, 64, 7)]


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.set_threads(8)
    return o


def run_sweeps(m, solid, f, wavefront):
    """Factor, FP64 and FP32 solves, lexicographic SOR and a fixed-iteration
    mixed-precision PCG-IC(0) on one context."""
    out = {}
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.set_option(lesgpu.OPT_WAVEFRONT, 1 if wavefront else 0)
        g.set_mask(solid)
        g.build_stencil()
        g.ic0_factor()
        out["L"] = g.get_stencil(lesgpu.STENCIL_L)
        out["ic0"] = g.precond_apply(lesgpu.PC_IC0, f[0])
        out["legacy"] = g.precond_apply(lesgpu.PC_LEGACY_IC_A, f[1])
        g.set_state({"u": f[0], "v": f[1], "w": f[2], "p": f[3] * 0.01})
        out["sweeps"] = g.sor_solve(1e-3, 1.2, 1.7, 7, lesgpu.SOLVER_SOR_LEX)
        out["p_sor"] = g.get_field("p")
        g.set_field("b", f[4])
        g.set_field("p", np.zeros(g.N))
        res = g.pcg(precond=lesgpu.PC_IC0, max_iter=12, tol=0.0, precision=lesgpu.PREC_MIXED,
                    reliable_every=5)
        out["iters"] = res["iters"]
        out["x_mixed"] = g.get_field("p")
    return out


@pytest.mark.parametrize("shape", SHAPES)
def test_wavefront_matches_hyperplanes(shape):
    m, solid, f = random_case(*shape, frac=0.05)
    a = run_sweeps(m, solid, f, True)
    b = run_sweeps(m, solid, f, False)
    for x, y in zip(a["L"], b["L"]):
        assert bits_equal(x, y)
    for key in ("ic0", "legacy", "p_sor", "x_mixed"):
        assert bits_equal(a[key], b[key]), key
    assert a["sweeps"] == b["sweeps"]
    assert a["iters"] == b["iters"]


@pytest.mark.parametrize("shape", [(64, 64, 64, 8), (72, 37, 50, 9)])
def test_wavefront_against_oracle(orc, shape):
    m, solid, f = random_case(*shape, frac=0.05)
    A = orc.build_stencil(solid, m)
    L_o = orc.ic_build(A, m)
    p0 = f[3] * 0.01
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.set_mask(solid)
        g.build_stencil()
        g.ic0_factor()
        for x, y in zip(g.get_stencil(lesgpu.STENCIL_L), L_o):
            assert bits_equal(x, y)
        assert bits_equal(g.precond_apply(lesgpu.PC_IC0, f[0]), orc.ic_apply(f[0], L_o, m))
        assert bits_equal(g.precond_apply(lesgpu.PC_LEGACY_IC_A, f[1]), orc.ic_apply(f[1], A, m))
        g.set_state({"u": f[0], "v": f[1], "w": f[2], "p": p0})
        s = g.sor_solve(1e-3, 1.2, 1.7, 5, lesgpu.SOLVER_SOR_LEX)
        po, so = orc.gs_solve(p0, f[0], f[1], f[2], solid, m, 1e-3, 1.2, 5)
        assert s == so
        assert bits_equal(g.get_field("p"), po)


def test_wavefront_repeated_launches(orc):
    """Many launches on one context (each with a new packet tag), results
    identical every time."""
    m, solid, f = random_case(40, 33, 21, 10, frac=0.05)
    A = orc.build_stencil(solid, m)
    L_o = orc.ic_build(A, m)
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.set_mask(solid)
        g.build_stencil()
        for rep in range(3):
            g.ic0_factor()
            for r in range(6):
                v = f[(rep + r) % 8]
                assert bits_equal(g.precond_apply(lesgpu.PC_IC0, v), orc.ic_apply(v, L_o, m)), (rep, r)


def test_wavefront_spin_limit():
    """A spin limit of 0 polls forbids every wait: the waves stop waiting,
    the launch completes and the call reports LES_ERR_STALL; with the limit
    restored the next calls are exact again."""
    m, solid, f = box_case(64, 11)
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.set_mask(solid)
        g.build_stencil()
        g.ic0_factor()
        ref = g.precond_apply(lesgpu.PC_IC0, f[0])
        with pytest.raises(lesgpu.LesError) as ei:
            g.set_option(lesgpu.OPT_SPIN_LIMIT, -1)
        assert ei.value.code == lesgpu.ERR_ARG
        g.set_option(lesgpu.OPT_SPIN_LIMIT, 0)
        stalls = 0
        for _ in range(3):
            try:
                z = g.precond_apply(lesgpu.PC_IC0, f[0])
            except lesgpu.LesError as e:
                assert e.code == lesgpu.ERR_STALL
                stalls += 1
            else:
                assert bits_equal(z, ref)
        # 64 tiles: a tile's first edge values are seven levels of another
        # wave away, so some wave always has to wait
        assert stalls > 0
        g.set_option(lesgpu.OPT_SPIN_LIMIT, 1 << 22)
        assert bits_equal(g.precond_apply(lesgpu.PC_IC0, f[0]), ref)
        g.ic0_factor()
        assert bits_equal(g.precond_apply(lesgpu.PC_IC0, f[0]), ref)
