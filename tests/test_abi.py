"""The built libraries load without a GPU and export every entry point their
headers declare: the C ABI (include/les_gpu.h, liblesgpu.so) and the C++
drop-in layer with the reference's signatures (include/les/pressure.h,
include/les/physics.h, liblespressure_gpu.so).  No compute call is made."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

import lesgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
DROPIN_SO = os.path.join(ROOT, "les-model_amd", "liblespressure_gpu.so")


def c_abi_names():
    text = open(os.path.join(INC, "les_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(les_\w+)\s*\(", text, re.M)))


def dropin_names():
    names = set()
    for h in ("pressure.h", "physics.h"):
        text = open(os.path.join(INC, "les", h)).read()
        names |= set(re.findall(r"^(?:void|int|double)\s+(\w+)\s*\(", text, re.M))
    return sorted(names)


def test_header_parse_finds_the_api():
    names = c_abi_names()
    for must in ("les_ctx_create", "les_pcg", "les_step", "les_rollout", "les_adjoint",
                 "les_last_error"):
        assert must in names
    assert len(dropin_names()) >= 25


def test_c_abi_exports_every_declared_symbol():
    lib = lesgpu.load_library()
    missing = [n for n in c_abi_names() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(shutil.which("nm") is None, reason="nm not available")
def test_dropin_exports_every_declared_function():
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", DROPIN_SO], check=True,
                         capture_output=True, text=True).stdout
    defined = set(re.findall(r"^\w+ [TW] (?:\w+::)*?(\w+)\(", out, re.M))
    missing = [n for n in dropin_names() if n not in defined]
    assert not missing, missing


def test_dropin_library_loads():
    C.CDLL(DROPIN_SO)
