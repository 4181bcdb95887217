"""The C-ABI libraries load and export every function their headers declare (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

from kcoregpu_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(k[a-z]{1,3}_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.mark.parametrize("header,lib", [("kcore.h", N.HIP_LIB), ("kcore_graph.h", N.HOST_LIB)])
def test_exports(header, lib):
    names = declared_functions(header)
    assert len(names) >= 8
    dll = ctypes.CDLL(lib)
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing


def test_status_strings():
    L = N.hip()
    for s in range(8):
        assert L.kc_status_str(s)
    assert L.kc_status_str(0) == b"ok"


def test_driver_usage_line():
    exe = os.path.join(N.BIN_DIR, "kcore")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.stdout == "Please provide data file\n"
    assert r.returncode != 0
