"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message (this build measures one GPU), and the algorithmic byte
counts it reports follow the kernels' read and write lists."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_pcg_bytes_per_iteration():
    import bench
    # mixed precision, coded operator, degree 8 with pairs: first pair 3e+2,
    # two later pairs 5e+2, the reducing step 4e+2, matvec 6e+2, update 3e
    assert bench.pcg_bytes_per_iter(8, 4, True, 2) == 14 + 2 * 22 + 18 + 26 + 12
    # single steps only: first 3e+op, five later 5e+op, the last 4e+op
    assert bench.pcg_bytes_per_iter(8, 8, False, 1) == (24 + 56) + 5 * (40 + 56) + (32 + 56) \
        + (48 + 56) + 24
    # the timed kernels' averages: first and two later pairs; the six
    # non-reducing single steps
    assert bench.cheb_bytes_per_cell(8, 4, True, 2) == (14 + 2 * 22) / 3
    assert bench.cheb_bytes_per_cell(8, 4, True, 1) == (14 + 5 * 22) / 6


def test_cheb_schedule_follows_the_chain():
    """The launch schedule mirrors the library's chaining: pairs never end
    on the reducing last step, single steps fill the rest."""
    import bench
    assert [w for w, _, _ in bench.cheb_schedule(8, 2)] == [2, 2, 2, 1]
    assert [w for w, _, _ in bench.cheb_schedule(7, 2)] == [2, 2, 1, 1]
    assert [w for w, _, _ in bench.cheb_schedule(4, 2)] == [2, 1]
    assert [w for w, _, _ in bench.cheb_schedule(3, 2)] == [1, 1]
    assert [w for w, _, _ in bench.cheb_schedule(8, 1)] == [1] * 7
