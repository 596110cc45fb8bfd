"""les_dot_noalloc (no reference counterpart): parallel_dot
(pressure.cpp:962) for callers that cannot allocate.  Its reduction
partials live in a static device array; the result must equal les_dot's on
the same grid bit for bit, on consecutive calls of different sizes (the
array keeps the previous call's contents) and on a caller's stream."""
import numpy as np
import pytest

import les_workloads as wl
import lesgpu

pytestmark = pytest.mark.gpu


def test_dot_noalloc_matches_les_dot():
    import torch
    torch.cuda.set_device(0)
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    stream = torch.cuda.Stream()   # default priority
    cases = ((64, 48, 40, 1, None), (9, 7, 13, 2, stream), (5, 3, 4, 3, None))
    for nx, ny, nz, seed, st in cases:
        N = nx * ny * nz
        print(f"[dot_noalloc] grid {nx}x{ny}x{nz}, seed {seed}, "
              f"{'caller stream' if st is not None else 'null stream'}", flush=True)
        a = torch.from_numpy(wl.uniform(seed, 0, N)).to("cuda")
        b = torch.from_numpy(wl.uniform(seed, 1, N)).to("cuda")
        out[:1].fill_(float("nan"))
        torch.cuda.synchronize()
        if st is None:
            lesgpu.dot_noalloc(nx, ny, nz, a, b, out)
        else:
            lesgpu.dot_noalloc(nx, ny, nz, a, b, out, stream=st.cuda_stream)
        torch.cuda.synchronize()
        got = out[:1].cpu().numpy()
        with lesgpu.LesGpu(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz) as g:
            ref = np.array([g.dot(a, b)])
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (nx, ny, nz, got, ref)
        # and the sequential sum in float64, to the summation-order bar
        seq = float(np.dot(a.cpu().numpy(), b.cpu().numpy()))
        assert abs(got[0] - seq) <= 1e-12 * max(1.0, abs(seq))
        assert float(out[1].item()) == 0.0   # nothing beyond *out is written


def test_dot_noalloc_rejects_grids_beyond_32bit_indices():
    import torch
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(lesgpu.LesError, match="2\\^31"):
        lesgpu.dot_noalloc(65536, 65536, 1, x, x, x)
    with pytest.raises(lesgpu.LesError, match=">= 1"):
        lesgpu.dot_noalloc(0, 4, 4, x, x, x)
