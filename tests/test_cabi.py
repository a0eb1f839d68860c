"""The C-ABI libraries load without a GPU and export every function that
include/*.h declares (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "constant-q-pstd_amd")

HEADERS = {"cq.h": "libcq_hip.so", "cq_host.h": "libcq_host.so", "cq_revolve.h": "libcq_host.so",
           "cq_lowrank.h": "libcq_host.so", "cq_shim.h": "libcq_shim.so"}


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(cq\w*)\s*\(", src, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_exports(header):
    path = os.path.join(PKG, HEADERS[header])
    if not os.path.exists(path):
        pytest.fail("%s not built" % path)
    lib = C.CDLL(path)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_without_gpu():
    # a fresh process: the error string is per thread and other CPU tests set
    # it through the same loaded library
    import subprocess
    import sys
    code = ("import ctypes as C; L = C.CDLL(%r); L.cq_version.restype = C.c_char_p; "
            "L.cq_last_error.restype = C.c_char_p; print(L.cq_version().decode(), repr(L.cq_last_error()))"
            % os.path.join(PKG, "libcq_hip.so"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True).stdout.split()
    assert out[0] == "cq-pstd-mi355x"
    assert out[-1] == "b''"
