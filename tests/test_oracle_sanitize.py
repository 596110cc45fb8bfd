"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/cpp/oracle_sanitize.c runs every part of oracle/les_oracle.c once on a
small anisotropic grid with obstacles (operator, preconditioners, every PCG
variant, the stationary solvers, physics, whole steps, the adjoint and the
statistics).  Built here with gcc -fsanitize=address,undefined and run on one
OpenMP thread; any out-of-bounds access, leak or undefined operation fails
the test.  The result must equal the plain build's bit for bit.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "oracle", "les_oracle.c"),
       os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c")]

needs_gcc = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")


def build(out, flags):
    cmd = ["gcc", "-std=c11", "-O1", "-g", "-fopenmp", "-ffp-contract=off",
           "-I" + os.path.join(ROOT, "oracle")] + flags + SRC + ["-o", out, "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def run(exe):
    env = dict(os.environ, OMP_NUM_THREADS="1", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    return subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)


@needs_gcc
def test_oracle_clean_under_asan_ubsan(tmp_path):
    plain, san = str(tmp_path / "plain"), str(tmp_path / "san")
    build(plain, [])
    build(san, ["-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                "-fno-sanitize-recover=undefined"])
    a, b = run(plain), run(san)
    assert a.returncode == 0, a.stderr[-2000:]
    assert b.returncode == 0, b.stderr[-4000:]
    assert "runtime error" not in b.stderr and "AddressSanitizer" not in b.stderr, b.stderr[-4000:]
    assert a.stdout.strip() == b.stdout.strip() and a.stdout.startswith("oracle checksum")
