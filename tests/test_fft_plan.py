"""FFT plans the engine takes (cq_fft_plan, include/cq.h): the host-side
planner needs no GPU.  Pinned by the plans of the workloads' lengths, by
properties every plan must have (the radices multiply to the length, are
supported radices, respect the per-stage capacity, end on a radix that is not
a composite), and by the stage count against the greedy factorisation that
the composite radices replace."""
import ctypes as C
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = C.CDLL(os.path.join(ROOT, "constant-q-pstd_amd", "libcq_hip.so"))
LIB.cq_fft_plan.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]

NT = 512                                   # threads per workgroup
NB = {2: 12, 3: 6, 4: 5, 5: 4, 7: 3, 8: 3, 9: 2, 11: 2, 15: 2}   # butterflies per thread, else 1
PRIMES = (2, 3, 5, 7, 11, 13, 17)
RADICES = {2, 3, 4, 5, 7, 8, 9, 11, 13, 16, 17, 15, 21, 25, 27}
COMPOSITES = {15, 21, 25, 27}


def plan(n, points, variant=0):
    rad = (C.c_int * 16)()
    inner = C.c_int(-1)
    ns = LIB.cq_fft_plan(n, points, variant, rad, 16, C.byref(inner))
    return (list(rad[:ns]) if ns >= 0 else None), inner.value


def smooth(n, primes=PRIMES):
    for p in primes:
        while n % p == 0:
            n //= p
    return n == 1


def greedy_stages(n, points):
    """Stages of the factorisation without composites: powers of two as
    16/8/4/2, then 9, 3, 5, 7, 11, 13, 17 (largest first, within capacity)."""
    m, k = n, 0
    for r in (16, 8, 4, 2, 9, 3, 5, 7, 11, 13, 17):
        if NT * NB.get(r, 1) * r < points:
            continue
        while m % r == 0:
            m //= r
            k += 1
    return k if m == 1 else None


@pytest.mark.parametrize("n,lines,want", [
    (680, 12, [8, 5, 17]),        # batched workload depth (24-column tiles)
    (1296, 6, [16, 9, 9]),        # batched workload interior width (12-depth tiles)
    (880, 6, [16, 5, 11]),        # Q-RTM workload depth
    (2100, 3, [25, 21, 4]),       # Q-RTM workload width (was 4 3 5 5 7)
    (4200, 2, [25, 21, 8]),       # largest serving width (was 8 3 5 5 7)
    (1050, 6, [25, 21, 2]),       # batched workload edge widths
    (750, 6, [25, 15, 2]),
    (1250, 6, [25, 25, 2]),
    (1080, 6, [8, 15, 9]),        # pad-period length: its power of two first
])
def test_workload_plans(n, lines, want):
    got, inner = plan(n, lines * n)
    assert inner == 0
    assert got == want


def test_variant1_small_radices():
    got, _ = plan(2100, 3 * 2100, 1)
    assert got and max(got) <= 8 or sorted(got) == sorted([4, 3, 5, 5, 7])
    for r in got:
        assert r in (2, 3, 4, 5, 7, 8, 11, 13, 17)
    assert all(r not in COMPOSITES for r in got)


def test_bluestein_inner_length():
    # 1281 = 3 7 61: Bluestein over the smallest 7-smooth M >= 2 n - 1
    got, inner = plan(1281, 2 * 1281)
    assert inner == 2592
    prod = 1
    for r in got:
        prod *= r
    assert prod == inner


@pytest.mark.parametrize("lines", [1, 2, 6, 12])
def test_plan_properties(lines):
    checked = 0
    for n in range(2, 4400):
        if not smooth(n):
            continue
        points = lines * n
        got, inner = plan(n, points)
        if got is None:
            assert greedy_stages(n, points) is None, n   # the greedy plan did not fit either
            continue
        assert inner == 0, n
        prod = 1
        for r in got:
            assert r in RADICES, (n, got)
            assert NT * NB.get(r, 1) * r >= points, (n, got)
            prod *= r
        assert prod == n, (n, got)
        assert got[-1] not in COMPOSITES, (n, got)
        g = greedy_stages(n, points)
        # never more stages than the greedy plan (which may not fit at all:
        # a composite stage holds more points than its factors' stages)
        assert g is None or len(got) <= g, (n, got, g)
        checked += 1
    assert checked > 100


def test_bad_arguments():
    # in a child process: the error string is per thread and would outlive
    # this test in the library the other CPU tests load
    import subprocess
    import sys
    code = ("import ctypes as C; L = C.CDLL(%r); r = (C.c_int * 4)(); "
            "print(L.cq_fft_plan(0, 8, 0, r, 4, None) < 0, L.cq_fft_plan(64, 32, 0, r, 4, None) < 0, "
            "L.cq_fft_plan(64, 64, 2, r, 4, None) < 0, L.cq_fft_plan(64, 64, 0, r, 4, None) > 0, "
            "L.cq_fft_plan(1, 1, 0, r, 4, None) == 0)"
            % os.path.join(ROOT, "constant-q-pstd_amd", "libcq_hip.so"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True).stdout
    assert out.split() == ["True"] * 5
