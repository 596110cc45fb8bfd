"""Checkpoint, VTK and statistics formats on the CPU.

The oracle's restatement of the reference's file formats
(oracle/les_io_oracle.py) and of its statistics arithmetic (les_oracle.c)
against the reference's own writers, built from src/io.cpp and
src/statistics.cpp into oracle/_ref/libref.so; and the library's host-side
VTK writer (les_write_field_vtk_host, no GPU) against both.  Cases:
negative zero, subnormals, values that round up a decade at 6 and 10
digits, huge and tiny exponents, mask bytes other than 0/1, and short or
missing checkpoint files.
"""
import os

import numpy as np
import pytest

import les_io_oracle as iox
from cpu_checkers import IO_KEYS, REF_SO, Oracle, Reference
from io_cases import LOAD_CASES, fields, load_case, mesh, write_ckpt

needs_ref = pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build absent")


@needs_ref
@pytest.mark.parametrize("shape", [(5, 4, 6), (3, 7, 3)])
def test_vtk_oracle_and_host_writer_match_reference(tmp_path, shape):
    import lesgpu
    m = mesh(*shape)
    f = fields(m, 1)
    ref = Reference().write_field_vtk(str(tmp_path), f, m, 12)
    want = open(ref, "rb").read()
    assert iox.vtk_fields_text(12, m, *[f[k] for k in IO_KEYS], f["solid"]) == want
    out = str(tmp_path / "lib.vtk")
    lesgpu.write_field_vtk_host(out, 12, m, *[f[k] for k in IO_KEYS], f["solid"])
    assert open(out, "rb").read() == want


@needs_ref
def test_vtk_host_writer_blocks_and_threads(tmp_path, monkeypatch):
    """More lines than one formatting block (2^20) on 7 threads, so block
    and thread boundaries fall mid-block; against the reference writer."""
    import lesgpu
    m = mesh(104, 100, 101)
    f = fields(m, 2)
    want = open(Reference().write_field_vtk(str(tmp_path), f, m, 300), "rb").read()
    monkeypatch.setenv("LES_IO_THREADS", "7")
    out = str(tmp_path / "lib.vtk")
    lesgpu.write_field_vtk_host(out, 300, m, *[f[k] for k in IO_KEYS], f["solid"])
    assert open(out, "rb").read() == want


def test_vtk_host_writer_reports_unwritable_path(tmp_path):
    import lesgpu
    m = mesh(3, 3, 3)
    f = fields(m, 3)
    with pytest.raises(lesgpu.LesError, match="cannot open"):
        lesgpu.write_field_vtk_host(str(tmp_path / "missing" / "x.vtk"), 0, m,
                                    *[f[k] for k in IO_KEYS], f["solid"])


@needs_ref
def test_checkpoint_format_oracle_matches_reference(tmp_path):
    m = mesh(6, 5, 4)
    f = fields(m, 4)
    for (t, dt, step) in ((0.1234567890123456789, 1e-3, 40), (2.0 / 3.0, 7.3e-5, 1000),
                          (0.0, 1.0, 0)):
        Reference().save_state(str(tmp_path), f, m, t, dt, step)
        d = tmp_path / "checkpoint"
        assert (d / "checkpoint.meta").read_bytes() == iox.checkpoint_meta_text(step, t, dt)
        assert (d / "checkpoint.data").read_bytes() == iox.checkpoint_data_bytes(
            *[f[k] for k in IO_KEYS], f["solid"])


@needs_ref
@pytest.mark.parametrize("case", LOAD_CASES)
def test_load_state_oracle_matches_reference(tmp_path, case):
    m = mesh(5, 4, 3)
    f_file, f_prev = fields(m, 5), fields(m, 6)
    meta, data = load_case(case, m, f_file)
    write_ckpt(tmp_path / "checkpoint", meta, data)
    step, t, dt, out = Reference().load_state(str(tmp_path), f_prev, m, 9.0, 8.0)
    o_step, o_t, o_dt, o_out = iox.load_state(str(tmp_path / "checkpoint"), f_prev, 9.0, 8.0)
    assert (step, t, dt) == (o_step, o_t, o_dt)
    for k in IO_KEYS:
        assert np.array_equal(out[k].view(np.uint8), o_out[k].view(np.uint8)), k
    # the reference's mask is a vector<bool>: compare as solid / not solid
    assert np.array_equal(out["solid"] != 0, o_out["solid"] != 0)


@needs_ref
def test_statistics_oracle_matches_reference(tmp_path):
    m = mesh(7, 6, 5)
    n = m["nx"] * m["ny"] * m["nz"]
    r = np.random.default_rng(7)
    steps = [998, 999, 1000, 1001, 1002, 1003, 1100]
    states = [{q: r.standard_normal(n) * 10.0 ** r.integers(-3, 3) for q in "uvwp"}
              for _ in steps]
    ns, ref_sums = Reference().statistics(str(tmp_path), m, 1000, steps, states)
    assert ns == 4
    orc = Oracle()
    sums = [np.zeros(n) for _ in range(10)]
    for s, st in zip(steps, states):
        if s > 1000:
            orc.stats_update(sums, st["u"], st["v"], st["w"], st["p"])
    for a, b in zip(sums, ref_sums):
        assert np.array_equal(a, b)
    out = orc.stats_finalize(sums, ns)
    want = (tmp_path / "output" / "statistics_final.vtk").read_bytes()
    assert iox.vtk_statistics_text(m, out) == want
