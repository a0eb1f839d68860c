"""Revolve schedule: the product's restatement vs the reference's src.c compiled
unmodified (oracle/_ref), plus the invariants of binomial checkpointing that
hold without the reference (numforw closed form, every step reversed once,
never more than snaps checkpoints)."""
import ctypes as C

import pytest

import refc
from cqpy import host

CASES = [(2501, 20), (3000, 40), (100, 5), (57, 3), (6000, 40), (10, 10), (33, 1), (2, 1),
         (1000, 64), (5, 20), (800, 7)]


def ref_trace(steps, snaps):
    L = refc.load("libref_revolve.so")
    check, capo, fine, info = C.c_int(-1), C.c_int(0), C.c_int(steps), C.c_int(3)
    out = []
    while True:
        a = L.revolve(C.byref(check), C.byref(capo), C.byref(fine), C.c_int(snaps),
                      C.byref(info))
        out.append((a, capo.value, check.value, fine.value))
        if a in (5, 6):
            return out, info.value


@pytest.mark.parametrize("steps,snaps", CASES)
def test_trace_matches_reference(steps, snaps):
    mine, info = host.revolve_trace(steps, snaps)
    ref, rinfo = ref_trace(steps, snaps)
    assert mine == ref
    assert info == rinfo
    L = refc.load("libref_revolve.so")
    assert host.numforw(steps, snaps) == L.numforw(C.c_int(steps), C.c_int(snaps))


@pytest.mark.parametrize("steps,snaps", [(2501, 20), (6000, 40), (100, 5), (57, 3), (50, 1)])
def test_schedule_invariants(steps, snaps):
    tr, info = host.revolve_trace(steps, snaps)
    assert tr[-1][0] == 5  # terminate
    reversed_steps, slots = [], set()
    for a, capo, check, fine in tr:
        assert check < snaps
        if a == 1:
            slots.add(check)
        if a in (3, 4):
            reversed_steps.append(capo)
    # every time step is reversed exactly once, from last to first
    assert reversed_steps == list(range(steps - 1, -1, -1))
    assert len(slots) <= snaps


def test_numforw_closed_form():
    # numforw(2501, 20) quoted in the survey of the reference: 7980 forward steps
    assert host.numforw(2501, 20) == 7980
    assert host.numforw(6000, 40) == 17097
    assert host.numforw(10, 0) == -1 and host.numforw(10, 65) == -1


def test_error_path_snaps_too_large():
    tr, info = host.revolve_trace(100, 65)
    assert tr[-1][0] == 6 and info in (11, 12, 14)
