"""GPU differentiation of the fixed-dt LES step map (les_rollout,
les_adjoint and the torch.autograd.Function over them).

* The GPU rollout and adjoint against the oracle's (orc_rollout,
  orc_adjoint, themselves pinned by central finite differences on the CPU),
  on an anisotropic grid with an obstacle in the mask (A built before the
  mask, as the shipped driver builds it), including an interior solid at
  i = nx-2 (the outlet-copy path), with checkpoint segments of unequal
  length.
* Recomputation is exact: gradients with checkpoints every step, every 5
  steps and only at step 0 are bit-identical.
* The differentiation workload: 96^3, FP64, 20 steps at dt = 1e-3 with
  checkpoints every 5 steps, loss = kinetic energy over a seeded 16^3
  sub-box at step 20, gradient with respect to the post-BC initial
  velocity through torch autograd, checked against central differences
  (eps = 1e-6) of the GPU forward on 64 seeded directions of unit RMS per
  cell to 1e-4 relative.
"""
import numpy as np
import pytest
import torch

import les_workloads as wl
import lesgpu
from cases import rel_err
from cpu_checkers import Oracle, oracle_rollout, oracle_adjoint, PC_JACOBI, PC_CHEBYSHEV

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    o = Oracle()
    o.set_threads(1)
    return o


def small_case(orc):
    m = dict(nx=12, ny=10, nz=9, dx=1.0 / 12, dy=0.9 / 10, dz=1.1 / 9)
    solid = wl.box_mask(12, 10, 9, [((4, 7), (3, 6), (3, 6))])
    solid[10 * 90 + 5 * 9 + 4] = 1          # interior solid at i = nx-2
    N = solid.size
    u0, v0, w0 = (wl.uniform(11, s, N) * 0.5 for s in range(3))
    u0, v0, w0 = orc.velocity_bcs(u0, v0, w0, solid, m)
    A = orc.build_stencil(np.zeros_like(solid), m)   # A before the obstacle exists
    return m, solid, (u0, v0, w0), A


@pytest.mark.parametrize("precond", [PC_JACOBI, PC_CHEBYSHEV])
def test_rollout_and_adjoint_match_oracle(orc, precond):
    m, solid, U0, A = small_case(orc)
    nsteps, every = 4, 3
    kw = dict(precond=precond, degree=4, tol_rel=1e-13)
    if precond == PC_CHEBYSHEV:
        kw.update(lmin=2.0, lmax=1.05 * orc.lambda_max(A, m, 40))
    uN, vN, wN, _, traj = oracle_rollout(orc, *U0, solid, m, A, nsteps, record=True, **kw)
    rng = np.random.default_rng(5)
    gbar = [rng.standard_normal(solid.size) for _ in range(3)]
    go = oracle_adjoint(orc, *gbar, traj, solid, m, A, nsteps, **kw)

    with lesgpu.LesGpu.from_mesh(m) as g:
        g.set_mask(solid)
        g.build_stencil(ignore_mask=True)
        for name, a in zip("uvw", U0):
            g.set_field(name, a)
        prm = g.adj_params(dt=1e-3, nsteps=nsteps, checkpoint_every=every, precond=precond,
                           degree=4, lmin=kw.get("lmin", 0.0), lmax=kw.get("lmax", 0.0))
        st = g.rollout(prm)
        assert st["solves"] == nsteps and st["unconverged"] == 0, st
        for name, ref in zip("uvw", (uN, vN, wN)):
            assert rel_err(g.get_field(name), ref) < 1e-10
        gg = [x.copy() for x in gbar]
        st = g.adjoint(prm, *gg)
        # recompute: 2 + 0 forward solves, then one adjoint solve per step
        assert st["solves"] == (every - 1) + nsteps and st["unconverged"] == 0, st
        for a, b in zip(gg, go):
            assert rel_err(a, b) < 1e-9, rel_err(a, b)
        # the adjoint leaves the rolled-out state in place
        assert rel_err(g.get_field("u"), uN) < 1e-10


def test_rejects_what_the_map_does_not_cover(orc):
    m, solid, U0, A = small_case(orc)
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.set_mask(solid)
        g.build_stencil()                       # A with mirror rows around the obstacle
        prm = g.adj_params(nsteps=2, checkpoint_every=1)
        with pytest.raises(lesgpu.LesError, match="obstacles"):
            g.rollout(prm)
        inlet = solid.copy()
        inlet[3 * 9 + 4] = 1                    # a solid on the inlet plane
        g.set_mask(inlet)
        g.build_stencil(ignore_mask=True)
        with pytest.raises(lesgpu.LesError, match="inlet"):
            g.rollout(prm)
        g.set_mask(solid)
        with pytest.raises(lesgpu.LesError, match="les_rollout"):
            g.adjoint(prm, *[np.zeros(solid.size) for _ in range(3)])


def test_checkpointing_is_exact(orc):
    m, solid, U0, A = small_case(orc)
    rng = np.random.default_rng(9)
    gbar = [rng.standard_normal(solid.size) for _ in range(3)]
    grads = []
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.set_mask(solid)
        g.build_stencil(ignore_mask=True)
        for every in (1, 5, 7):
            for name, a in zip("uvw", U0):
                g.set_field(name, a)
            prm = g.adj_params(nsteps=7, checkpoint_every=every)
            g.rollout(prm)
            gg = [x.copy() for x in gbar]
            g.adjoint(prm, *gg)
            grads.append(np.concatenate(gg))
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def test_autograd_gradient_vs_finite_differences_96():
    """The differentiation workload at its full size."""
    from les_autograd import rollout

    n, nsteps, every, side = 96, 20, 5, 16
    m = wl.mesh(n)
    N = n ** 3
    dv = m["dx"] * m["dy"] * m["dz"]
    dev = torch.device("cuda:0")
    solid = np.zeros(N, dtype=np.uint8)
    rng = np.random.default_rng(2024)
    lo = rng.integers(1, n - side - 1, size=3)
    box = np.zeros((n, n, n))
    box[lo[0]:lo[0] + side, lo[1]:lo[1] + side, lo[2]:lo[2] + side] = 1.0
    box_t = torch.from_numpy(box.reshape(-1)).to(dev)

    g = lesgpu.LesGpu.from_mesh(m)
    g.set_mask(solid)
    g.build_stencil()
    u, v, w = wl.spectral_velocity(n, 0)
    for name, a in zip("uvw", (u, v, w)):
        g.set_field(name, np.ascontiguousarray(a.reshape(-1)))
    g.velocity_bcs()                          # the post-BC initial state
    U0 = [torch.from_numpy(g.get_field(name)).to(dev) for name in "uvw"]
    prm = g.adj_params(dt=1e-3, nsteps=nsteps, checkpoint_every=every, Cs=0.17)

    def loss_of(uN, vN, wN):
        return 0.5 * torch.sum(box_t * (uN * uN + vN * vN + wN * wN)) * dv

    X = [t.clone().requires_grad_(True) for t in U0]
    L = loss_of(*rollout(*X, g, prm))
    L.backward()
    grad = [x.grad.detach() for x in X]
    assert all(torch.isfinite(t).all() for t in grad)

    # Central differences of the GPU forward (the same rollout, pressure
    # solved to 1e-13 relative) along seeded directions of unit RMS per cell:
    # each velocity value moves by about eps, so the differences resolve the
    # map far above the solves' tolerance.
    eps = 1e-6
    drng = np.random.default_rng(64)
    errs = []
    with torch.no_grad():
        for _ in range(64):
            d = [torch.from_numpy(drng.standard_normal(N)).to(dev) for _ in range(3)]
            lp = loss_of(*rollout(*[a + eps * b for a, b in zip(U0, d)], g, prm)).item()
            lm = loss_of(*rollout(*[a - eps * b for a, b in zip(U0, d)], g, prm)).item()
            fd = (lp - lm) / (2 * eps)
            ad = sum(torch.sum(a * b) for a, b in zip(grad, d)).item()
            errs.append(abs(fd - ad) / abs(fd))
    print(f"[adjoint 96^3] loss {L.item():.6e}; FD vs adjoint rel err over 64 directions: "
          f"max {max(errs):.2e}, median {float(np.median(errs)):.2e}", flush=True)
    g.close()
    assert max(errs) <= 1e-4, errs
