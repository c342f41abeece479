"""The C-ABI library loads and exports every function include/gexs_hip.h declares.

No compute call is made here (no GPU in the CPU suite); only calls that are
rejected on their arguments before touching the device.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

from gexs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "gexs_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.findall(r"^\s*(?:int|const char\s*\*)\s*(gexs_\w+)\s*\(", src, flags=re.M)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        subprocess.check_call(["python", "-m", "gexs_amd.build"], cwd=ROOT)
    return _lib.load()


def test_header_and_binding_agree():
    assert declared() == _lib.EXPORTS


def test_every_declared_symbol_is_exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (gexs_\w+)$", out, flags=re.M))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing


def test_null_context_is_rejected(lib):
    assert lib.gexs_destroy(None) == -1
    assert lib.gexs_sync(None) == -1
    assert lib.gexs_ready(None) == -1
    assert lib.gexs_last_error(None) == b"null context"
    h = C.c_void_p()
    assert lib.gexs_create(0, 3, C.byref(h)) == -1  # PSUEDO_VOIGT is not a profile here
    assert lib.gexs_create(0, 2, None) == -1
