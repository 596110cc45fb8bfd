"""Shared inputs of the checkpoint / VTK / statistics tests: fields with
values that stress number formatting, and the checkpoint files of each
load_state case (whole, missing, short in a field or in the mask, odd mask
bytes, trailing bytes)."""
import numpy as np

from cpu_checkers import IO_KEYS
import les_io_oracle as iox


SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, 0.1, 1e-310, -2.5e-320, 999999.5, 9999995.0,
                    0.000123456789, 1.5e300, -7.25e-300, 123456.5, 1234565.0, 99999.95,
                    9.9999999995, 3.0, 1e22, 2.0 ** -1074, 1.7976931348623157e308])


def mesh(nx, ny, nz):
    return dict(nx=nx, ny=ny, nz=nz, dx=1.0 / (nx - 1), dy=0.2, dz=1.0 / 3.0)


def fields(m, seed):
    n = m["nx"] * m["ny"] * m["nz"]
    r = np.random.default_rng(seed)
    f = {}
    for q, k in enumerate(IO_KEYS):
        a = r.standard_normal(n) * 10.0 ** r.integers(-12, 12, n)
        a[: len(SPECIAL)] = np.roll(SPECIAL, q)[: min(n, len(SPECIAL))]
        f[k] = a
    f["solid"] = r.integers(0, 3, n).astype(np.uint8)  # 2 counts as solid
    return f


def write_ckpt(d, meta, data):
    d.mkdir(parents=True, exist_ok=True)
    for name, content in (("checkpoint.meta", meta), ("checkpoint.data", data)):
        p = d / name
        if p.exists():
            p.unlink()
        if content is not None:
            p.write_bytes(content)


LOAD_CASES = ["whole", "no_meta", "no_data", "short_in_v", "short_in_mask", "mask_empty",
              "empty_data", "mask_bytes", "long_data"]


def load_case(case, m, f_file):
    n = m["nx"] * m["ny"] * m["nz"]
    meta = iox.checkpoint_meta_text(77, 0.375, 2.5e-4)
    data = iox.checkpoint_data_bytes(*[f_file[k] for k in IO_KEYS], f_file["solid"])
    if case == "no_meta":
        meta = None
    elif case == "no_data":
        data = None
    elif case == "short_in_v":
        data = data[: 8 * n + 8 * (n // 3) + 3]
    elif case == "short_in_mask":
        data = data[: 40 * n + n // 2]
    elif case == "mask_empty":
        data = data[: 40 * n]
    elif case == "empty_data":
        data = b""
    elif case == "mask_bytes":
        data = data[: 40 * n] + bytes([0, 1, 2, 255] * n)[:n]
    elif case == "long_data":
        data = data + b"\x01" * 17
    return meta, data
