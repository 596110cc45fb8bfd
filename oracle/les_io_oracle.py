"""CPU restatement of the reference's file formats and checkpoint reader.

TEST INFRASTRUCTURE ONLY: the tests use it as the checker of the library's
checkpoint, VTK and statistics writers; the product path never imports it.

* VTK (io.cpp:225-256) and statistics VTK (statistics.cpp:11-67): ASCII
  text; doubles go through std::ostream's default floatfield, which is
  printf's %.<precision>g in the "C" locale: precision 6, except the
  statistics data after its header (std::setprecision(10),
  statistics.cpp:34).  Python's '%.*g' is the same correctly rounded
  conversion with the same layout rules (two-digit exponent minimum,
  trailing zeros stripped).
* checkpoint.meta (io.cpp:20-30): the step as %d, then time and dt with
  precision max_digits10 = 17.
* checkpoint.data (io.cpp:33-57): u, v, w, p, nu_t as native doubles, then
  one byte per cell, 1 for solid and 0 otherwise.
* load_state (io.cpp:61-120): restated below with the stream semantics the
  reference relies on (a short read fills a prefix and fails the stream;
  the mask buffer starts zeroed and becomes byte == 1).

Pinned against the reference build (oracle/_ref/libref.so) by
tests/test_io_cpu.py.
"""
from __future__ import annotations

import os

import numpy as np


def fmt_g(x: float, prec: int = 6) -> str:
    return "%.*g" % (prec, x)


def _header(title: str, m) -> str:
    return ("# vtk DataFile Version 3.0\n%s\nASCII\nDATASET STRUCTURED_POINTS\n"
            "DIMENSIONS %d %d %d\nORIGIN 0 0 0\nSPACING %s %s %s\nPOINT_DATA %d\n"
            % (title, m["nx"], m["ny"], m["nz"], fmt_g(m["dx"]), fmt_g(m["dy"]), fmt_g(m["dz"]),
               m["nx"] * m["ny"] * m["nz"]))


def _scalars(name: str, a, prec: int) -> str:
    body = "".join(fmt_g(x, prec) + "\n" for x in np.asarray(a, dtype=np.float64).tolist())
    return "SCALARS %s double 1\nLOOKUP_TABLE default\n%s" % (name, body)


def _vectors(name: str, u, v, w, prec: int) -> str:
    rows = zip(*(np.asarray(a, dtype=np.float64).tolist() for a in (u, v, w)))
    body = "".join("%s %s %s\n" % (fmt_g(a, prec), fmt_g(b, prec), fmt_g(c, prec))
                   for a, b, c in rows)
    return "VECTORS %s double\n%s" % (name, body)


def vtk_fields_text(step: int, m, u, v, w, p, nu_t, solid) -> bytes:
    """write_field_vtk (io.cpp:225-256)."""
    sol = "".join("1\n" if s else "0\n" for s in np.asarray(solid).tolist())
    t = (_header("LES output step %d" % step, m) + _vectors("velocity", u, v, w, 6)
         + _scalars("pressure", p, 6) + _scalars("nu_t", nu_t, 6)
         + "SCALARS solid int 1\nLOOKUP_TABLE default\n" + sol)
    return t.encode()


def vtk_statistics_text(m, out10) -> bytes:
    """write_statistics_vtk (statistics.cpp:11-67); out10 = means of u, v,
    w, p, then the stresses uu, vv, ww, uv, uw, vw."""
    t = _header("Time-averaged statistics", m) + _vectors("mean_velocity", *out10[:3], 10)
    t += _scalars("mean_pressure", out10[3], 10)
    for name, a in zip(("uu", "vv", "ww", "uv", "uw", "vw"), out10[4:]):
        t += _scalars(name + "_stress", a, 10)
    return t.encode()


def checkpoint_meta_text(next_step: int, time_simulated: float, dt: float) -> bytes:
    """checkpoint.meta (io.cpp:20-30)."""
    return ("%d\n%s\n%s\n" % (next_step, fmt_g(time_simulated, 17), fmt_g(dt, 17))).encode()


def checkpoint_data_bytes(u, v, w, p, nu_t, solid) -> bytes:
    """checkpoint.data (io.cpp:33-57)."""
    parts = [np.ascontiguousarray(a, dtype=np.float64).tobytes() for a in (u, v, w, p, nu_t)]
    parts.append((np.asarray(solid) != 0).astype(np.uint8).tobytes())
    return b"".join(parts)


def load_state(directory: str, fields: dict, time_simulated: float, dt: float):
    """load_state (io.cpp:61-120) on copies of fields (u, v, w, p, nu_t
    float64 and solid uint8 of one length n).  Returns (step, time, dt,
    fields)."""
    out = {k: np.array(v, copy=True) for k, v in fields.items()}
    meta = os.path.join(directory, "checkpoint.meta")
    if not os.path.exists(meta):
        return 0, time_simulated, dt, out
    tok = open(meta, "rb").read().split()
    # operator>> on int then two doubles: a failed extraction stores 0 and
    # fails the stream, so later extractions leave their targets unchanged
    vals, ok = [], True
    for q, conv in enumerate((int, float, float)):
        if ok and q < len(tok):
            try:
                vals.append(conv(tok[q]))
                continue
            except ValueError:
                ok = False
                vals.append(0)
                continue
        if ok:  # end of file: extraction fails, 0 stored
            ok = False
            vals.append(0)
        else:
            vals.append(None)
    start = vals[0]
    if vals[1] is not None:
        time_simulated = vals[1]
    if vals[2] is not None:
        dt = vals[2]
    data = os.path.join(directory, "checkpoint.data")
    if not os.path.exists(data):
        return 0, time_simulated, dt, out
    raw = open(data, "rb").read()
    n = len(out["u"])
    pos, failed = 0, False
    for k in ("u", "v", "w", "p", "nu_t"):
        want = 8 * n
        got = 0 if failed else min(want, len(raw) - pos)
        if got:
            buf = bytearray(out[k].tobytes())
            buf[:got] = raw[pos:pos + got]
            out[k] = np.frombuffer(bytes(buf), dtype=np.float64).copy()
        pos += got
        failed = failed or got < want
    got = 0 if failed else min(n, len(raw) - pos)
    mask = np.zeros(n, dtype=np.uint8)
    mask[:got] = np.frombuffer(raw[pos:pos + got], dtype=np.uint8) == 1
    out["solid"] = mask
    if got == 0:
        return 0, time_simulated, dt, out
    return start, time_simulated, dt, out
