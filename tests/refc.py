"""Loader for oracle/_ref: the reference's own host C files, compiled
unmodified from the reference tree by oracle/Makefile (target ``ref``).
Used only as a checker by the CPU tests; absent libraries skip those tests."""
import ctypes as C
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def load(name):
    path = os.path.join(REF_DIR, name)
    if not os.path.exists(path):
        pytest.skip("oracle/_ref/%s not built (reference tree absent)" % name)
    return C.CDLL(path)
