#!/usr/bin/env python3
"""LES steps/s at 256^3 on one MI355X: FP64 vs mixed-precision PCG to 1e-8.

Workload (the mixed-precision configuration of the reference's LES step):
  256^3 unit cube, spectral velocity field (E(k) ~ k^-5/3, rms 1, seed 0),
  5% of cells solid (Bernoulli, seed 0), p = 0, nu = 5e-3, rho = 1.2,
  Cs = 0.17.  The configuration leaves nu open; at nu = 1e-3 the
  reference's forward-Euler step with central advection (main.cpp:563-659)
  is unstable on this field: dt = 1e-4 exceeds the FTCS limit
  2 nu_eff / |u|^2 ~ 8e-5, and max |u| grows about 14% a step from step
  ~27 (nu = 2e-3 likewise); at nu = 5e-3 it stays bounded for 100 steps.
  The pressure operator A is built before the obstacle mask, as
  the shipped driver builds it (main.cpp:65-81): built from the mask, its
  identity rows hold p = 0 inside every solid cell next to mirror-Neumann
  fluid rows, correct_velocity (main.cpp:671-714) differences p across that
  jump, and the reference's own step map diverges within a few steps.
  Steps are numbered from 21, so the reference's adaptive time step
  (main.cpp:115-119) sets dt: the fixed dt = 1e-3 of steps 0-20 puts the
  explicit scheme at CFL ~1.3 at 256^3.  Pressure by PCG-Chebyshev degree 8
  on bounds from 20 Lanczos steps, to ||r|| <= 1e-8 ||b||; mixed precision
  replaces the residual in FP64 every 50 iterations.
One step = Smagorinsky nu_t -> momentum (+ lagged divergence/RHS) -> PCG ->
projection, all on the GPU with fields resident in HBM.

Every step's PCG iterations, dt and max |u| are recorded; a mode whose
trajectory fails (a solve that does not converge, a non-finite scalar, or
max |u| growing past GROWTH_LIMIT times its first value) is refused: its
timing is dropped and its entry carries the error.  The JSON line is printed
in every case; the exit status is non-zero when the headline mode failed.

`value` is the fastest mixed-precision mode's steps/s; every mode's line is
reported beside it.  Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "les-model_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# the Chebyshev launch of each fusion width and its mode suffix
# the preconditioner's timed launch per mode suffix: Chebyshev as single
# steps (1) or fused pairs (2), or the multigrid V-cycle's level-0 sweeps (0)
KERNEL_OF_WIDTH = {0: "k_mg_smooth", 1: "k_cheb_step", 2: "k_cheb_pair"}
WIDTH_NAME = {0: "mg", 1: "single", 2: "pairs"}
METRIC = "LES steps/s at 256^3, FP64 vs mixed PCG to 1e-8 rel.; PCG HBM GB/s vs peak"
GROWTH_LIMIT = 4.0   # max |u| may grow to this multiple of its value at the first step


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cheb_schedule(degree, width):
    """The launches of one Chebyshev application as the library chains them
    (les_api.hip cheb_chain): fused launches of `width` steps (1 single, 2
    pairs) that never run the reducing last step, then single steps.
    (steps, first, last) each."""
    out, k = [], 1
    while k < degree:
        w = width
        while w > 1 and (degree - k < w or k + w - 1 == degree - 1):
            w -= 1
        out.append((w, k == 1, k + w - 1 == degree - 1))
        k += w
    return out


def _launch_bytes(first, last, elem, coded):
    # first: r, the operator -> p, z.  Later: z, r, p, the operator -> p, z
    # (a fused launch's intermediate z and p stay on chip).  The reducing
    # last step stores z only.
    op = 2 if coded else 7 * elem
    if last:
        return 4 * elem + op
    return (3 if first else 5) * elem + op


def cheb_bytes_per_cell(degree, elem, coded=False, width=1):
    """Algorithmic bytes per cell per timed Chebyshev launch of the fused
    width (1: k_cheb_step, 2: k_cheb_pair), averaged
    over that width's launches of one application that the library times
    (all but a reducing last one).  The operator is 7 coefficient arrays,
    or one 2-byte code.  Width 0: the multigrid's level-0 pre-smoothing,
    its two damped-Jacobi sweeps from x = 0 in one pass (b, the operator ->
    x), in FP32 with the seven arrays with the restriction in the same pass
    (and b of the next level, an eighth of the cells)."""
    if width == 0:
        return 2 * elem + (2 if coded else 7 * elem) + (elem / 8 if elem == 4 and not coded else 0)
    sched = cheb_schedule(degree, width)
    timed = [_launch_bytes(f, l, elem, coded) for w, f, l in sched if w == width and not l]
    return sum(timed) / len(timed)


def pcg_bytes_per_iter(degree, elem, coded, width):
    """Algorithmic bytes per cell of one PCG-Chebyshev iteration as the
    library runs it: the preconditioner application (fused launches of the
    given width where they apply, the reducing last step single), the matvec
    with the direction update and the previous iteration's x update fused in
    (p_old, z, the operator, x -> p, q, x) and the r update (r, q -> r).
    The finalize kernels and the gated FP64 residual replacement (every 50
    iterations) are left out."""
    op = 2 if coded else 7 * elem
    cheb = sum(_launch_bytes(f, l, elem, coded) for _, f, l in cheb_schedule(degree, width))
    return cheb + (6 * elem + op) + 3 * elem


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "les_step_256_traffic.json")


def measured_traffic(mode, n):
    """HBM bytes per launch of the Chebyshev step kernel in this mode, from
    the committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, each
    scaled by what the counter reads for the plain copy kernel), or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("n") != n:
        return None
    m = rec.get("modes", {}).get(mode)
    return None if m is None else m.get("traffic_bytes_per_launch")


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def make_workload(n, seed=0):
    import les_workloads as wl
    log(f"generating {n}^3 spectral field and 5% mask")
    u, v, w = wl.spectral_velocity(n, seed)
    solid = wl.bernoulli_mask(n, 0.05, seed)
    return wl.mesh(n), (u, v, w), solid


def step_record(s, r):
    return {"step": s, "iters": int(r["iters"]), "dt": r["dt"], "speed_max": r["speed_max"],
            "converged": int(r["converged"]), "fault": int(r["fault"]),
            "true_rel_res": r["true_res_norm"] / r["b_norm"] if r["b_norm"] > 0 else 0.0}


def trajectory_error(rec, speed0):
    """Why a step's record refuses the trajectory, or None."""
    if rec["fault"]:
        return f"step {rec['step']}: non-finite PCG scalar"
    if not rec["converged"]:
        return f"step {rec['step']}: PCG did not converge in {rec['iters']} iterations"
    sm = rec["speed_max"]
    if sm >= 0 and not math.isfinite(sm):
        return f"step {rec['step']}: max |u| is not finite"
    if sm >= 0 and speed0 > 0 and sm > GROWTH_LIMIT * speed0:
        return (f"step {rec['step']}: max |u| {sm:.3g} exceeds {GROWTH_LIMIT:g}x its first "
                f"value {speed0:.3g} (diverging)")
    return None


def run_mode(g, st0, precision, stencil, bounds, steps, warmup, kt_class, pcg_kw, step0,
             nu, keep_start_state=False, precond=None):
    import torch
    import lesgpu
    g.set_state(st0)
    lmin, lmax = bounds
    precond = lesgpu.PC_CHEBYSHEV if precond is None else precond
    # multigrid runs in the mirror-weighted inner product (CG on W A, which is
    # symmetric for the operator built before the mask)
    if precond == lesgpu.PC_MG:
        pcg_kw = dict(pcg_kw, inner_product=lesgpu.IP_MIRROR)
    kw = dict(nu=nu, rho=1.2, Cs=0.17, precond=precond, lmin=lmin, lmax=lmax,
              degree=8, max_iter=20000, tol=1e-8, tol_rel=True, precision=precision,
              reliable_every=50, stencil_format=stencil, **pcg_kw)
    traj = []
    speed0 = None
    for s in range(step0, step0 + warmup):
        rec = step_record(s, g.step(step=s, **kw))
        traj.append(rec)
        speed0 = rec["speed_max"] if speed0 is None else speed0
        log(f"  warmup step {s}: iters {rec['iters']} dt {rec['dt']:.3g} "
            f"max|u| {rec['speed_max']:.3f}")
        err = trajectory_error(rec, speed0)
        if err:
            return dict(error=err, trajectory=traj)
    start_state = g.get_state() if keep_start_state else None
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = []
    for s in range(step0 + warmup, step0 + warmup + steps):
        out.append((s, g.step(step=s, **kw)))
    g.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s, r in out:
        rec = step_record(s, r)
        traj.append(rec)
        speed0 = rec["speed_max"] if speed0 is None else speed0
        err = trajectory_error(rec, speed0)
        if err:
            return dict(error=err, trajectory=traj)
    state = g.get_state()
    last_speed = float(max(np.max(np.abs(state[k])) for k in ("u", "v", "w")))
    # the kernel's own duration, from HIP events around each of its launches
    # in one more step (the events are kept out of the timed steps: each
    # record adds a few microseconds between launches)
    g.kernel_timing(kt_class)
    g.step(step=step0 + warmup + steps, **kw)
    kt = g.kernel_timing_read()
    g.kernel_timing(lesgpu.KT_NONE)
    its = [rec["iters"] for rec in traj[warmup:]]
    return dict(seconds=t1 - t0, iters=its, trajectory=traj, kt=kt, state=state,
                start_state=start_state, final_component_max=last_speed)


def cpu_baseline(m, state, solid, bounds, threads, step, gpu_iters, nu, single_thread=False,
                 full_step=False):
    """The reference solver built from its own sources (oracle/_ref), else
    the oracle, timed on the host cores on the step the GPU timed first:
    from the GPU's FP64 state at that step, with the same stencil order,
    Chebyshev bounds and tolerance (std::chrono around the step inside the
    reference harness).

    A whole step takes about a minute on 16 threads, so the default is a
    bounded sample: the same step run twice, stopped after N1 and N2 PCG
    iterations.  The difference is the time of N2 - N1 iterations, the rest
    is the step outside the solve (physics, set-up, projection); the two are
    scaled to the iteration count the step takes to 1e-8 (the FP64 GPU
    solve's: the reference's count equals it, see the C2 replay fixtures).
    full_step times the whole step to 1e-8 instead."""
    from cpu_checkers import Reference, Oracle, REF_SO, PC_CHEBYSHEV
    lmin, lmax = bounds
    kw = dict(lmin=lmin, lmax=lmax, degree=8, tol=1e-8, tol_rel=True, nu=nu, rho=1.2,
              max_iter=20000)
    if os.path.exists(REF_SO):
        chk, kind = Reference(), "reference"
    else:
        chk, kind = Oracle(), "port"
    # the shipped driver's order: A built before the mask (main.cpp:65-81)
    A = chk.build_stencil(np.zeros_like(solid), m)

    def one_step(nthreads, max_iter=None):
        chk.set_threads(nthreads)
        secs = []
        t0 = time.perf_counter()
        kws = dict(kw, max_iter=max_iter) if max_iter else kw
        if kind == "reference":
            # the reference fixes Cs = 0.18 (main.cpp:425); the cost is the same
            _, its, dts = chk.run_steps(state, solid, m, A, 1, step0=step, solver=4,
                                        step_seconds=secs, **kws)
        else:
            _, its, dts = chk.run_steps(state, solid, m, A, 1, step0=step, precond=PC_CHEBYSHEV,
                                        Cs=0.17, **kws)
            secs = [time.perf_counter() - t0]
        return float(secs[0]), int(its[0]), float(dts[0])

    def sampled(nthreads, n1, n2, iters):
        ta, _, _ = one_step(nthreads, n1)
        tb, _, dt = one_step(nthreads, n2)
        per_it = (tb - ta) / (n2 - n1)
        other = max(ta - n1 * per_it, 0.0)
        return other + iters * per_it, per_it, other, ta, tb, dt

    common = dict(unit="steps/s", kind=kind, nproc=os.cpu_count(), cpu_model=cpu_model(),
                  gpu_pcg_iters=gpu_iters)
    if full_step:
        t, it, dt = one_step(threads)
        out = dict(common, value=1.0 / t, cores=threads, seconds_per_step=t, pcg_iters=it,
                   pcg_iters_per_s=it / t,
                   sample=f"step {step} of the {m['nx']}^3 workload (the first step the GPU "
                          f"timed) from the GPU's FP64 state at that step: the reference's "
                          f"solve_pressure_pcg_chebyshev(degree 8) on the same Lanczos bounds "
                          f"to 1e-8 relative, A built before the mask; {it} PCG iterations "
                          f"(GPU: {gpu_iters}), dt {dt:.3g}, {t:.1f} s on {threads} OpenMP "
                          f"threads of {cpu_model()}")
    else:
        n1, n2 = 20, 100
        t, per_it, other, ta, tb, dt = sampled(threads, n1, n2, gpu_iters)
        out = dict(common, value=1.0 / t, cores=threads, seconds_per_step=t,
                   pcg_iters=gpu_iters, pcg_iters_per_s=1.0 / per_it,
                   seconds_per_pcg_iter=per_it, seconds_outside_pcg=other,
                   sample=f"step {step} of the {m['nx']}^3 workload (the first step the GPU "
                          f"timed) from the GPU's FP64 state at that step, the reference's "
                          f"solve_pressure_pcg_chebyshev(degree 8) on the same Lanczos bounds, "
                          f"A built before the mask, run twice on {threads} OpenMP threads of "
                          f"{cpu_model()} and stopped after {n1} and {n2} PCG iterations "
                          f"({ta:.1f} s and {tb:.1f} s): {per_it:.4f} s per iteration and "
                          f"{other:.2f} s outside the solve, scaled to the {gpu_iters} "
                          f"iterations the step takes to 1e-8 relative")
    if single_thread:
        # the same step on one thread, stopped after 20 and 60 iterations
        n1, n2 = 20, 60
        log(f"CPU baseline, one thread: the same step stopped after {n1} and {n2} iterations")
        t1, per_it, other, ta, tb, _ = sampled(1, n1, n2, out["pcg_iters"])
        out["single_thread"] = dict(
            value=1.0 / t1, unit="steps/s", cores=1, seconds_per_step=t1,
            pcg_iters=out["pcg_iters"], seconds_per_pcg_iter=per_it, seconds_outside_pcg=other,
            sample=f"the same step on one thread stopped after {n1} and {n2} PCG iterations "
                   f"({ta:.1f} s and {tb:.1f} s): {per_it:.3f} s per iteration and {other:.1f} s "
                   f"outside the solve, scaled to the step's {out['pcg_iters']} iterations")
    chk.set_threads(threads)
    return out


def adjoint_timing():
    """The differentiation workload (96^3, FP64, 20 fixed-dt steps with
    checkpoints every 5, pressure to 1e-13 relative): wall time of the
    rollout and of the adjoint (which recomputes each segment)."""
    import lesgpu
    import les_workloads as wl
    n = 96
    m = wl.mesh(n)
    g = lesgpu.LesGpu.from_mesh(m)
    g.set_mask(np.zeros(n ** 3, dtype=np.uint8))
    u, v, w = wl.spectral_velocity(n, 0)
    for name, a in zip("uvw", (u, v, w)):
        g.set_field(name, a)
    g.velocity_bcs()
    g.build_stencil()
    prm = g.adj_params(dt=1e-3, nsteps=20, checkpoint_every=5, Cs=0.17)
    g.synchronize()
    t0 = time.perf_counter()
    fwd = g.rollout(prm)
    t1 = time.perf_counter()
    ub = [np.ones(n ** 3) * m["dx"] ** 3 for _ in range(3)]
    bwd = g.adjoint(prm, *ub)
    t2 = time.perf_counter()
    g.close()
    return {"workload": "96^3 FP64, 20 steps dt=1e-3, checkpoints every 5, Jacobi-PCG in the "
                        "mirror-weighted inner product to 1e-13 rel",
            "rollout_s": t1 - t0, "adjoint_s": t2 - t1, "rollout_pcg_iters": fwd["iters"],
            "adjoint_pcg_iters": bwd["iters"], "unconverged": fwd["unconverged"] +
            bwd["unconverged"]}


def sweep_timing(g, m, hyperplanes=True):
    """The lexicographic sweeps on the workload's grid and operator: the IC(0)
    factor (pressure.cpp:495-556), its forward + backward solve in FP64 and
    in FP32 (pressure.cpp:442-482, inside a fixed-iteration mixed PCG-IC(0))
    and one lexicographic SOR sweep (pressure.cpp:216-234), each as one
    wavefront launch and, for comparison, as one launch per i+j+k
    hyperplane.  Times are HIP events around the launches of the class.
    Algorithmic bytes per cell: factor 4 reads + 4 writes, solve pair 2 x (5
    reads + 1 write), SOR sweep p and rhs read + p written (neighbour
    values are re-reads of p)."""
    import lesgpu
    N = m["nx"] * m["ny"] * m["nz"]
    levels = m["nx"] + m["ny"] + m["nz"] - 2
    r = np.random.default_rng(5).uniform(-1.0, 1.0, N)
    st = g.get_state()

    def timed(cls, fn, reps):
        g.kernel_timing(cls)
        for _ in range(reps):
            fn()
        n_l, ms, _ = g.kernel_timing_read()
        g.kernel_timing(lesgpu.KT_NONE)
        return 1e3 * ms / max(n_l, 1)

    def mixed_ic():
        g.set_field("p", np.zeros(N))
        g.pcg(precond=lesgpu.PC_IC0, max_iter=6, tol=0.0, precision=lesgpu.PREC_MIXED,
              reliable_every=50)

    def sor():
        g.set_field("p", np.zeros(N))
        g.sor_solve(1e-3, 1.0, 1.7, 2, lesgpu.SOLVER_SOR_LEX)

    g.set_field("b", r)
    cases = (("ic0_factor", lesgpu.KT_IC_FACTOR, g.ic0_factor, 8 * 8, "k_wf<double, FACTOR>"),
             ("ic0_solve_fp64", lesgpu.KT_IC_SOLVE,
              lambda: g.precond_apply(lesgpu.PC_IC0, r), 2 * 6 * 8,
              "k_wf<double, FWD> + k_wf<double, BWD>"),
             ("ic0_solve_fp32", lesgpu.KT_IC_SOLVE, mixed_ic, 2 * 6 * 4,
              "k_wf<float, FWD> + k_wf<float, BWD>"),
             ("sor_lex_sweep", lesgpu.KT_SOR_SWEEP, sor, 3 * 8, "k_wf<double, SOR>"))
    out = {"levels": levels}
    g.ic0_factor()
    for wf in ((1, 0) if hyperplanes else (1,)):
        g.set_option(lesgpu.OPT_WAVEFRONT, wf)
        for name, cls, fn, bpc, kern in cases:
            fn()  # warm
            us = timed(cls, fn, 3 if wf else 1)
            e = out.setdefault(name, {"kernel": kern, "bytes_per_cell": bpc})
            if wf:
                gbs = bpc * N / (us * 1e-6) / 1e9
                e.update(avg_us=us, achieved_gbs=gbs, frac=gbs / HBM_PEAK_GBS,
                         ns_per_level=1e3 * us / levels / (2 if "solve" in name else 1))
            else:
                e["hyperplane_launches_us"] = us
                e["speedup_vs_hyperplanes"] = us / e["avg_us"]
    g.set_option(lesgpu.OPT_WAVEFRONT, 1)
    g.set_state(st)
    return out


def io_timing(g, m, ref_sample=True):
    """The next-row paths on the workload's grid, from the fields the timed
    run left on the device: the statistics update kernel (HIP events around
    each launch), the checkpoint save and load, and the VTK writer (blocking
    and background), with files written under TMPDIR and removed.  The
    reference's VTK writer is timed on a 64^3 sample and scaled by the cell
    count (bounded CPU work)."""
    import shutil
    import tempfile
    import lesgpu
    N = m["nx"] * m["ny"] * m["nz"]
    out = {}
    g.stats_init(0)
    for s in range(2):
        g.stats_update(1 + s)
    g.kernel_timing(lesgpu.KT_STATS)
    reps = 20
    for s in range(reps):
        g.stats_update(3 + s)
    n_l, ms, _ = g.kernel_timing_read()
    g.kernel_timing(lesgpu.KT_NONE)
    us = 1e3 * ms / max(n_l, 1)
    gbs = 192.0 * N / (us * 1e-6) / 1e9
    out["stats_update"] = {"kernel": "k_stats_update", "avg_us": us, "launches": n_l,
                           "bytes_per_cell": 192, "achieved_gbs": gbs,
                           "frac": gbs / HBM_PEAK_GBS}
    d = tempfile.mkdtemp(prefix="les_io_", dir=os.environ.get("TMPDIR"))
    try:
        g.synchronize()
        t0 = time.perf_counter()
        g.save_state(1.0, 1e-3, 100, d)
        t1 = time.perf_counter()
        g.load_state(d)
        t2 = time.perf_counter()
        size = os.path.getsize(os.path.join(d, "checkpoint.data"))
        out["checkpoint"] = {"bytes": size, "save_s": t1 - t0, "load_s": t2 - t1,
                             "save_gbs": size / (t1 - t0) / 1e9,
                             "load_gbs": size / (t2 - t1) / 1e9}
        os.remove(os.path.join(d, "checkpoint.data"))
        path = os.path.join(d, "out.vtk")
        t0 = time.perf_counter()
        g.write_field_vtk(100, path)
        t1 = time.perf_counter()
        vsize = os.path.getsize(path)
        os.remove(path)
        g.synchronize()
        t2 = time.perf_counter()
        g.write_field_vtk(200, path, background=True)
        t3 = time.perf_counter()
        g.io_wait()
        t4 = time.perf_counter()
        out["vtk"] = {"bytes": vsize, "threads": int(os.environ.get("LES_IO_THREADS", "0")) or
                      min(16, os.cpu_count() or 1), "write_s": t1 - t0,
                      "background_call_s": t3 - t2, "background_total_s": t4 - t2}
        if ref_sample:
            from cpu_checkers import REF_SO, Reference
            if os.path.exists(REF_SO):
                import les_workloads as wl
                n = 64
                ms_ = wl.mesh(n)
                r = np.random.default_rng(0)
                f = {k: r.standard_normal(n ** 3) for k in ("u", "v", "w", "p", "nu_t")}
                f["solid"] = np.zeros(n ** 3, np.uint8)
                t0 = time.perf_counter()
                Reference().write_field_vtk(d, f, ms_, 0)
                t1 = time.perf_counter()
                out["vtk"]["reference_writer_s_scaled"] = (t1 - t0) * N / n ** 3
                out["vtk"]["reference_sample"] = "write_field_vtk (io.cpp:225) on 64^3, 1 thread"
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def pcg_hbm(r, elem, cells):
    """Whole-iteration PCG bandwidth of one mode: algorithmic bytes per
    iteration times iterations per second of the timed steps (their physics
    counted in the time), against the 8 TB/s peak."""
    if r["width"] == 0:
        return None   # the V-cycle's bytes are not modelled
    bpc = pcg_bytes_per_iter(8, elem, r["coded"], r["width"])
    gbs = bpc * cells * float(np.sum(r["iters"])) / r["seconds"] / 1e9
    return {"bytes_per_cell_per_iter": bpc, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS}


def parse_mode(mode):
    parts = mode.split(":")
    prec_s = parts[0]
    sten_s = parts[1] if len(parts) > 1 else "coded"
    kern_s = parts[2] if len(parts) > 2 else "pairs"
    return prec_s, sten_s, {"mg": 0, "single": 1, "pairs": 2}[kern_s]


def base_line(args, m=None):
    grid = [args.n] * 3 if m is None else [m["nx"], m["ny"], m["nz"]]
    return {
        "metric": METRIC, "value": None, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded spectral velocity, seeded 5% solid mask)",
        "config": {"workload": f"LES step {grid[0]}^3, spectral velocity, 5% Bernoulli solids, "
                               f"nu {args.nu:g}, "
                               f"A built before the mask (main.cpp:65-81), PCG-Chebyshev(8) on "
                               f"20-step Lanczos bounds to 1e-8 rel, steps {args.step0 + args.warmup}"
                               f"-{args.step0 + args.warmup + args.steps - 1}",
                   "grid": grid, "solid_fraction": 0.05, "nu": args.nu, "rho": 1.2, "Cs": 0.17,
                   "precond": "chebyshev", "degree": 8,
                   "reliable_every": 50, "first_step": args.step0,
                   "stencil_order": "A before the mask (main.cpp:65-81)",
                   "dt": "adaptive (main.cpp:208-250)"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nu", type=float, default=5e-3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1))
    ap.add_argument("--cpu-single-thread", action="store_true",
                    help="also time the CPU baseline step on one thread (about 1.5 min at 256^3)")
    ap.add_argument("--cpu-full-step", action="store_true",
                    help="time the CPU baseline's whole step to 1e-8 (about a minute on 16 "
                         "threads at 256^3) instead of the bounded two-run sample")
    ap.add_argument("--modes", default="fp64:arrays,mixed:arrays,fp64:coded,mixed:coded,"
                                       "fp64:arrays:mg,mixed:coded:mg",
                    help="precision:stencil[:kind] entries; precision fp64|mixed, stencil "
                         "arrays|coded, kind single|pairs (default): PCG-Chebyshev(8) with its "
                         "steps single or fused in pairs, mg: PCG with one multigrid V-cycle "
                         "(reported beside the headline, which is PCG-Chebyshev)")
    ap.add_argument("--check-every", type=int, default=0)
    ap.add_argument("--step0", type=int, default=21)
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the timing of the differentiation workload")
    ap.add_argument("--no-sweeps", action="store_true",
                    help="skip the timing of the IC(0) and lexicographic SOR sweeps")
    ap.add_argument("--no-sweep-hyperplanes", action="store_true",
                    help="time the sweeps' wavefront launches only, not also one launch per "
                         "hyperplane (about 29000 small launches: slow under a profiler)")
    ap.add_argument("--no-io", action="store_true",
                    help="skip the statistics / checkpoint / VTK timings")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="launch the plain FP64 copy kernel this many times first "
                         "(reference size for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    out = base_line(args)
    try:
        rc = measure(args, out)
    except Exception as e:  # reported in the line, never swallowed silently
        import traceback
        traceback.print_exc()
        out["error"] = repr(e)
        rc = 1
    print(json.dumps(out), flush=True)
    sys.exit(rc)


def measure(args, out):
    import torch
    import lesgpu
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)

    m, (u, v, w), solid = make_workload(args.n)
    out.update(base_line(args, m))
    N = m["nx"] * m["ny"] * m["nz"]
    g = lesgpu.LesGpu.from_mesh(m)
    g.set_mask(solid)
    st0 = {k: np.zeros(N) for k in lesgpu.FIELDS[:9]}
    st0["u"], st0["v"], st0["w"] = u, v, w
    g.set_state(st0)
    g.velocity_bcs()          # main.cpp:102
    st0 = g.get_state()
    g.build_stencil(ignore_mask=True)   # main.cpp:65-81: A before the mask
    if args.copy_calibration:
        g.diag_copy(args.copy_calibration)
        log(f"copy calibration: {args.copy_calibration} x k_convert<double,double> of "
            f"{8 * N} bytes read + {8 * N} written")
    t0 = time.perf_counter()
    rmin, rmax = g.lanczos_bounds(20, seed=0)
    bounds = (rmin, 1.05 * rmax)
    log(f"Lanczos(20): ritz [{rmin:.4g}, {rmax:.4g}] -> Chebyshev [{bounds[0]:.4g}, "
        f"{bounds[1]:.4g}] in {time.perf_counter() - t0:.2f} s")
    out["config"]["lanczos_bounds"] = list(bounds)

    modes = args.modes.split(",")
    KT_OF_WIDTH = {0: lesgpu.KT_MG_SWEEP, 1: lesgpu.KT_CHEB_STEP, 2: lesgpu.KT_CHEB_PAIR}
    res = {}
    pcg_kw = dict(check_every=args.check_every)
    fp64_for_cpu = next((md for md in modes if md.startswith("fp64")), None)
    for mode in modes:
        prec_s, sten_s, width = parse_mode(mode)
        prec = lesgpu.PREC_MIXED if prec_s == "mixed" else lesgpu.PREC_FP64
        sten = lesgpu.STENCIL_CODED if sten_s == "coded" else lesgpu.STENCIL_ARRAYS
        g.set_option(lesgpu.OPT_CHEB_PAIRS, 1 if width != 1 else 0)
        log(f"mode {mode}: {args.warmup} warmup + {args.steps} timed steps")
        r = run_mode(g, st0, prec, sten, bounds, args.steps, args.warmup, KT_OF_WIDTH[width],
                     pcg_kw, args.step0, args.nu,
                     keep_start_state=(mode == fp64_for_cpu and not args.no_cpu_baseline),
                     precond=lesgpu.PC_MG if width == 0 else lesgpu.PC_CHEBYSHEV)
        r.update(elem=4 if prec_s == "mixed" else 8, coded=sten_s == "coded", width=width)
        res[mode] = r
        if "error" in r:
            log(f"  {mode}: REFUSED: {r['error']}")
            continue
        n_l, ms, _ = r["kt"]
        r["cheb_bpc"] = cheb_bytes_per_cell(8, r["elem"], r["coded"], width)
        r["cheb_gbs"] = (r["cheb_bpc"] * N) / (ms / n_l * 1e-3) / 1e9 if n_l else 0.0
        log(f"  {mode}: {args.steps / r['seconds']:.3f} steps/s, iters/step "
            f"{np.mean(r['iters']):.1f}, {KERNEL_OF_WIDTH[width]} {n_l} launches avg "
            f"{1e3 * ms / max(n_l, 1):.1f} us = {r['cheb_gbs']:.0f} GB/s algorithmic")

    def mode_entry(r):
        e = {"trajectory": r["trajectory"]}
        if "error" in r:
            e["error"] = r["error"]
            return e
        e.update(steps_per_s=args.steps / r["seconds"],
                 pcg_iters_per_step=float(np.mean(r["iters"])),
                 pcg_hbm=pcg_hbm(r, r["elem"], N),
                 cheb_kernel=KERNEL_OF_WIDTH[r["width"]],
                 cheb_gbs=r["cheb_gbs"], cheb_frac=r["cheb_gbs"] / HBM_PEAK_GBS,
                 cheb_bytes_per_cell=r["cheb_bpc"],
                 final_component_max=r["final_component_max"])
        return e

    out["modes"] = {mname: mode_entry(r) for mname, r in res.items()}
    # the headline is the configured method, PCG-Chebyshev(8): the multigrid
    # modes are reported beside it
    ok = [md for md in res if "error" not in res[md] and res[md]["width"] != 0]
    mixed_ok = [md for md in ok if md.startswith("mixed")]
    pool = mixed_ok or ok
    head = max(pool, key=lambda md: args.steps / res[md]["seconds"]) if pool else None
    head_wanted = "mixed" if any(md.startswith("mixed") for md in modes) else modes[0]
    rc = 0
    if head is None or not head.startswith(head_wanted.split(":")[0]):
        out["error"] = "the headline (mixed-precision) mode failed: " + "; ".join(
            f"{md}: {res[md]['error']}" for md in res if "error" in res[md])
        rc = 1
    if head is not None:
        R = res[head]
        steps_per_s = args.steps / R["seconds"]
        n_launch, kt_ms, elem = R["kt"]
        bpc = R["cheb_bpc"]
        achieved = R["cheb_gbs"]
        out.update({
            "value": steps_per_s,
            "ms_per_step": 1e3 * R["seconds"] / args.steps,
            "dtype": "f32+f64" if head.startswith("mixed") else "f64",
            "pcg_iters_per_step": float(np.mean(R["iters"])),
            "pcg_iters_per_s": float(np.sum(R["iters"]) / R["seconds"]),
            "roofline": {"bound": "hbm",
                         "kernel": f"{KERNEL_OF_WIDTH[R['width']]}"
                                   f"<{'float' if elem == 4 else 'double'}>"
                                   f"{' coded' if R['coded'] else ''}",
                         "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(
                             f"{head}:{WIDTH_NAME[R['width']]}", m["nx"]),
                         "algorithmic_bytes_per_launch": bpc * N,
                         "bytes_per_cell": bpc, "cells": N, "launches": n_launch,
                         "avg_us": 1e3 * kt_ms / n_launch if n_launch else None},
            "pcg_hbm": pcg_hbm(R, elem, N),
        })
        out["config"]["mode"] = head
        out["config"]["workload"] += f", {head}"
        rf = out["roofline"]
        if rf["traffic"]:
            # counter bytes below the algorithmic count would mean the
            # Infinity Cache served compulsory traffic: flagged
            rf["traffic_over_algorithmic"] = rf["traffic"] / rf["algorithmic_bytes_per_launch"]
            rf["traffic_below_algorithmic"] = rf["traffic_over_algorithmic"] < 1.0
        fp64_plain = "fp64:arrays" if "fp64:arrays" in ok else next(
            (md for md in ok if md.startswith("fp64")), None)
        if fp64_plain is not None:
            F = res[fp64_plain]
            out["fp64_plain_mode"] = fp64_plain
            out["fp64_steps_per_s"] = args.steps / F["seconds"]
            if head.startswith("mixed"):
                out["mixed_vs_fp64_speedup"] = steps_per_s / out["fp64_steps_per_s"]
                err = {}
                for k in ("u", "v", "w", "p"):
                    a, b = R["state"][k], F["state"][k]
                    err[k] = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
                out["mixed_max_rel_err_vs_fp64"] = err
    if not args.no_adjoint:
        log("differentiation workload: 96^3 rollout + adjoint")
        out["adjoint"] = adjoint_timing()
        log(f"  rollout {out['adjoint']['rollout_s']:.2f} s, adjoint "
            f"{out['adjoint']['adjoint_s']:.2f} s")
    if not args.no_sweeps:
        log("lexicographic sweeps (IC(0), SOR) on the workload's grid")
        out["lexicographic_sweeps"] = sweep_timing(g, m, not args.no_sweep_hyperplanes)
        log(f"  {json.dumps(out['lexicographic_sweeps'])}")
    if not args.no_io:
        log("statistics, checkpoint and VTK output on the workload's grid")
        out["io"] = io_timing(g, m, ref_sample=not args.no_cpu_baseline)
        log(f"  {json.dumps(out['io'])}")
    g.close()
    F = res.get(fp64_for_cpu) if fp64_for_cpu else None
    if not args.no_cpu_baseline and F is not None and F.get("start_state") is not None:
        step = args.step0 + args.warmup
        log(f"CPU baseline: step {step} of the reference on {args.cpu_threads} threads")
        try:
            out["cpu_baseline"] = cpu_baseline(m, F["start_state"], solid, bounds,
                                               args.cpu_threads, step, F["iters"][0],
                                               args.nu, single_thread=args.cpu_single_thread,
                                               full_step=args.cpu_full_step)
            log(f"  {out['cpu_baseline']['sample']}")
        except Exception as e:  # reported, not hidden
            out["cpu_baseline"] = {"error": repr(e)}
    return rc


if __name__ == "__main__":
    main()
