"""Checkpoint, VTK output and time statistics of the device-resident state.

The library's les_save_state / les_load_state / les_write_field_vtk(_async)
/ les_stats_* against the oracle's restatement of the reference's formats
and statistics arithmetic (oracle/les_io_oracle.py, les_oracle.c) and, where
the reference build is present, against the reference's own writers.  Bar:
byte-identical files and bit-identical fields and sums.  A resumed run
equals the continuous one bit for bit (no solid cell on the inlet plane,
the reference's own condition for that, main.cpp:340 / io.cpp:42-46).
"""
import os

import numpy as np
import pytest

import les_workloads as wl
import lesgpu

import les_io_oracle as iox
from cpu_checkers import IO_KEYS, REF_SO, Oracle, Reference
from io_cases import LOAD_CASES, fields, load_case, mesh, write_ckpt

pytestmark = pytest.mark.gpu
have_ref = os.path.exists(REF_SO)


def ctx_with(m, f):
    g = lesgpu.LesGpu.from_mesh(m)
    for k in IO_KEYS:
        g.set_field(k, f[k])
    g.set_mask(f["solid"])
    return g


def dev_fields(g):
    out = {k: g.get_field(k) for k in IO_KEYS}
    out["solid"] = g.get_mask()
    return out


@pytest.mark.parametrize("shape", [(6, 5, 4), (7, 5, 3), (40, 33, 29)])
def test_save_state_bytes(tmp_path, shape):
    m = mesh(*shape)
    f = fields(m, 11)
    with ctx_with(m, f) as g:
        g.save_state(0.3125 + 1e-17, 4.0e-4 / 3.0, 123, str(tmp_path))
    meta = (tmp_path / "checkpoint.meta").read_bytes()
    data = (tmp_path / "checkpoint.data").read_bytes()
    assert meta == iox.checkpoint_meta_text(123, 0.3125 + 1e-17, 4.0e-4 / 3.0)
    assert data == iox.checkpoint_data_bytes(*[f[k] for k in IO_KEYS], f["solid"])
    if have_ref:
        Reference().save_state(str(tmp_path / "ref"), f, m, 0.3125 + 1e-17, 4.0e-4 / 3.0, 123)
        assert (tmp_path / "ref" / "checkpoint" / "checkpoint.data").read_bytes() == data
        assert (tmp_path / "ref" / "checkpoint" / "checkpoint.meta").read_bytes() == meta


@pytest.mark.parametrize("case", LOAD_CASES)
def test_load_state_cases(tmp_path, case):
    m = mesh(5, 4, 3)
    f_file, f_prev = fields(m, 5), fields(m, 6)
    meta, data = load_case(case, m, f_file)
    write_ckpt(tmp_path / "checkpoint", meta, data)
    want = iox.load_state(str(tmp_path / "checkpoint"), f_prev, 9.0, 8.0)
    with ctx_with(m, f_prev) as g:
        step, t, dt = g.load_state(str(tmp_path / "checkpoint"), 9.0, 8.0)
        got = dev_fields(g)
    assert (step, t, dt) == want[:3], case
    for k in IO_KEYS + ("solid",):
        assert np.array_equal(got[k].view(np.uint8), want[3][k].view(np.uint8)), (case, k)
    if have_ref:
        r = Reference().load_state(str(tmp_path), f_prev, m, 9.0, 8.0)
        assert r[:3] == (step, t, dt)


def test_checkpoint_round_trip_128(tmp_path):
    """Save and load at 128^3 (more than one 32 MiB staging chunk per field)."""
    n = 128
    m = wl.mesh(n)
    N = n ** 3
    r = np.random.default_rng(3)
    f = {k: r.standard_normal(N) for k in IO_KEYS}
    f["solid"] = wl.random_boxes_mask(n, 3, 2)
    with ctx_with(m, f) as g:
        g.save_state(1.5, 2e-4, 50, str(tmp_path))
    assert os.path.getsize(tmp_path / "checkpoint.data") == 41 * N
    with ctx_with(m, {k: np.zeros(N) for k in IO_KEYS} | {"solid": np.zeros(N, np.uint8)}) as g:
        assert g.load_state(str(tmp_path)) == (50, 1.5, 2e-4)
        got = dev_fields(g)
    for k in IO_KEYS + ("solid",):
        assert np.array_equal(got[k], f[k]), k


def run(g, steps, st):
    for s in steps:
        r = g.step(step=s, nu=1e-3, rho=1.2, Cs=0.18, precond=lesgpu.PC_JACOBI, max_iter=5000,
                   tol=1e-9, tol_rel=True)
        st["time"] += r["dt"]
        st["dt"] = r["dt"]


def test_resume_equals_continuous(tmp_path):
    """main.cpp's loop with a checkpoint at step 22 (after the fixed-dt
    steps, so the adaptive dt is exercised on both sides) and a resume in a
    fresh context, against the run that never stopped."""
    m = mesh(24, 20, 16)
    n = 24 * 20 * 16
    solid = np.zeros((24, 20, 16), np.uint8)
    solid[8:12, 6:10, 4:9] = 1  # clear of the inlet plane i = 0
    solid = solid.ravel()
    r = np.random.default_rng(9)
    init = {k: 0.3 * r.standard_normal(n) for k in ("u", "v", "w")}

    def fresh():
        g = lesgpu.LesGpu.from_mesh(m)
        g.set_mask(solid)
        g.set_state(dict(init))
        g.velocity_bcs()
        g.build_stencil(ignore_mask=True)  # main.cpp:66 builds A before the mask
        return g

    cont = {"time": 0.0, "dt": 1e-3}
    with fresh() as g:
        run(g, range(19, 25), cont)
        want = dev_fields(g)
    part = {"time": 0.0, "dt": 1e-3}
    with fresh() as g:
        run(g, range(19, 22), part)
        g.save_state(part["time"], part["dt"], 22, str(tmp_path))
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.build_stencil(ignore_mask=True)
        step, t, dt = g.load_state(str(tmp_path))
        assert (step, t, dt) == (22, part["time"], part["dt"])
        part["time"] = t
        run(g, range(step, 25), part)
        got = dev_fields(g)
    assert part["time"] == cont["time"]
    for k in IO_KEYS + ("solid",):
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("shape", [(5, 4, 6), (7, 5, 3)])
def test_write_field_vtk_bytes(tmp_path, shape):
    m = mesh(*shape)
    f = fields(m, 12)
    want = iox.vtk_fields_text(400, m, *[f[k] for k in IO_KEYS], f["solid"])
    with ctx_with(m, f) as g:
        g.write_field_vtk(400, str(tmp_path / "a.vtk"))
    assert (tmp_path / "a.vtk").read_bytes() == want
    if have_ref:
        ref = Reference().write_field_vtk(str(tmp_path), f, m, 400)
        assert open(ref, "rb").read() == want


def test_write_field_vtk_async_snapshot(tmp_path):
    """The background writer sees the fields as they were at the call: a
    step queued right after it changes them on the device."""
    m = mesh(20, 18, 16)
    n = 20 * 18 * 16
    f = fields(m, 13)
    for k in ("u", "v", "w"):
        f[k] = np.random.default_rng(1).standard_normal(n)
    f["p"] = np.zeros(n)
    f["nu_t"] = np.zeros(n)
    f["solid"] = np.zeros(n, np.uint8)
    want = iox.vtk_fields_text(7, m, *[f[k] for k in IO_KEYS], f["solid"])
    with ctx_with(m, f) as g:
        g.build_stencil()
        g.write_field_vtk(7, str(tmp_path / "a.vtk"), background=True)
        g.step(step=3, nu=1e-3, rho=1.2, Cs=0.18, precond=lesgpu.PC_JACOBI, max_iter=50,
               tol=1e-6)
        g.write_field_vtk(8, str(tmp_path / "b.vtk"), background=True)  # waits for the first
        g.io_wait()
        after = dev_fields(g)
    assert (tmp_path / "a.vtk").read_bytes() == want
    assert not np.array_equal(after["u"], f["u"])
    assert (tmp_path / "b.vtk").read_bytes() == iox.vtk_fields_text(
        8, m, *[after[k] for k in IO_KEYS], after["solid"])


def test_io_wait_reports_failed_background_write(tmp_path):
    m = mesh(4, 4, 4)
    f = fields(m, 14)
    with ctx_with(m, f) as g:
        g.write_field_vtk(0, str(tmp_path / "missing" / "a.vtk"), background=True)
        with pytest.raises(lesgpu.LesError, match="cannot open"):
            g.io_wait()
        g.io_wait()  # the failure is reported once


@pytest.mark.parametrize("shape", [(7, 5, 3), (12, 10, 8)])
def test_statistics(tmp_path, shape):
    m = mesh(*shape)
    n = m["nx"] * m["ny"] * m["nz"]
    r = np.random.default_rng(21)
    steps = [998, 999, 1000, 1001, 1002, 1003, 1100]
    states = [{q: r.standard_normal(n) * 10.0 ** r.integers(-3, 3) for q in "uvwp"}
              for _ in steps]
    orc = Oracle()
    sums = [np.zeros(n) for _ in range(10)]
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.stats_init(1000)
        for s, st in zip(steps, states):
            for q in "uvwp":
                g.set_field(q, st[q])
            g.stats_update(s)
            if s > 1000:
                orc.stats_update(sums, st["u"], st["v"], st["w"], st["p"])
        assert g.stats_samples() == 4
        for q, name in enumerate(lesgpu.STATS):
            assert np.array_equal(g.stats_get(name), sums[q]), name
        g.stats_finalize(str(tmp_path / "stats.vtk"))
        out = orc.stats_finalize(sums, 4)
        for q, name in enumerate(lesgpu.STATS):
            assert np.array_equal(g.stats_get(name, final=True), out[q]), name
    text = (tmp_path / "stats.vtk").read_bytes()
    assert text == iox.vtk_statistics_text(m, out)
    if have_ref:
        Reference().statistics(str(tmp_path / "ref"), m, 1000, steps, states)
        assert (tmp_path / "ref" / "output" / "statistics_final.vtk").read_bytes() == text


def test_statistics_without_samples(tmp_path):
    m = mesh(4, 4, 4)
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.stats_init(1000)
        g.stats_update(1000)  # not after the start step
        g.stats_finalize(str(tmp_path / "s.vtk"))
        assert g.stats_samples() == 0
    assert not (tmp_path / "s.vtk").exists()  # statistics.cpp:135-138


def test_statistics_256_matches_numpy():
    """At the workload's size: the sums over three samples equal numpy's
    same-order additions bit for bit (each is one rounding per sample)."""
    n = 256
    N = n ** 3
    m = wl.mesh(n)
    r = np.random.default_rng(5)
    with lesgpu.LesGpu.from_mesh(m) as g:
        g.stats_init(0)
        acc = {"u": np.zeros(N), "uv": np.zeros(N), "p": np.zeros(N)}
        for s in (1, 2, 3):
            st = {q: r.standard_normal(N) for q in "uvwp"}
            for q in "uvwp":
                g.set_field(q, st[q])
            g.stats_update(s)
            acc["u"] += st["u"]
            acc["uv"] += st["u"] * st["v"]
            acc["p"] += st["p"]
        for k in acc:
            assert np.array_equal(g.stats_get(k), acc[k]), k
