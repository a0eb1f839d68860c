"""Failure paths through the C ABI: a shot that blows up (time step beyond the
CFL limit) is reported as CQ_ENUMERIC instead of as a result, a job whose
admission fails leaves the server's queue as failed while the other jobs
finish unchanged, and the plain path refuses RTM shots.  Every kernel still
runs to its end: a blown-up wavefield only holds inf/NaN values."""
import os

import numpy as np
import pytest

from cqpy import engine, synth
from test_gpu_rtm import rtm_job

pytestmark = pytest.mark.gpu

MAXX, MAXZ = 72, 48
ENUMERIC, ENOMEM, EINVAL = -6, -3, -1


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


def unstable_job(nt=600):
    """A small acoustic shot with dt five times its stable value (CFL ~2)."""
    j = synth.small_job(nxt=40, nzt=24, pml=5, nt=nt, kind=0, seed=31)
    j = dict(j)
    j["dt"] = 5 * j["dt"]
    return j


def code_of(exc):
    return int(str(exc).split("cq error ", 1)[1].split(":", 1)[0])


def run_alone(ctx, job):
    s = ctx.server(MAXX, MAXZ)
    i = s.submit(job)
    while s.state(i) not in (engine.DONE, engine.FAILED):
        s.advance(64)
    out = s.result(i)
    s.close()
    return out


def test_batch_reports_blown_up_shot(ctx):
    good = synth.small_job(nxt=40, nzt=24, pml=5, nt=600, kind=0, seed=32)
    b = ctx.batch([good, unstable_job()])
    b.run()
    rec = b.rec(0)
    assert np.isfinite(rec).all()
    with pytest.raises(engine.CqError) as e:
        b.rec(1)
    assert code_of(e.value) == ENUMERIC
    b.close()


def test_plain_path_stops_at_blow_up_and_refuses_rtm(ctx):
    with pytest.raises(engine.CqError) as e:
        ctx.plain_forward(unstable_job(nt=300))
    assert code_of(e.value) == ENUMERIC
    with pytest.raises(engine.CqError) as e:
        ctx.plain_forward(rtm_job(2, nt=20, snaps=4, seed=33))
    assert code_of(e.value) == EINVAL


@pytest.mark.parametrize("run_cells", [0, 1])
def test_server_blown_up_job_fails_others_unchanged(ctx, run_cells):
    """run_cells = 1: one job advances at a time (shortest remaining work
    first), the blown-up job among them."""
    good = [synth.small_job(nxt=40, nzt=24, pml=5, nt=700, kind=0, seed=34),
            synth.small_job(nxt=41, nzt=25, pml=7, nt=650, kind=1, seed=35)]
    alone = [run_alone(ctx, j) for j in good]
    s = ctx.server(MAXX, MAXZ, run_cells=run_cells)
    ids = [s.submit(good[0]), s.submit(unstable_job(nt=2000)), s.submit(good[1])]
    while any(s.state(i) in (engine.QUEUED, engine.RUNNING) for i in ids):
        s.advance(100)
    assert s.state(ids[1]) == engine.FAILED
    # the periodic check retires it long before its 2000 steps (all jobs
    # advancing at once; with one at a time it may wait for the others first)
    assert s.steps(ids[1])[2] < (2000 if run_cells == 0 else 2000 + 700 + 650)
    with pytest.raises(engine.CqError) as e:
        s.result(ids[1])
    assert code_of(e.value) == ENUMERIC
    for k, i in enumerate((ids[0], ids[2])):
        assert s.state(i) == engine.DONE
        assert np.array_equal(s.result(i), alone[k]), k
    s.release(ids[1])
    s.close()


def test_server_failed_admission_others_finish(ctx, monkeypatch):
    js = [synth.small_job(nxt=40, nzt=24, pml=5, nt=120, kind=0, seed=36),
          rtm_job(3, nt=60, snaps=5, seed=37),
          synth.small_job(nxt=52, nzt=20, pml=6, nt=90, kind=0, seed=38)]
    alone = [run_alone(ctx, j) for j in js]
    monkeypatch.setenv("CQ_FAULT_ADMIT", "2")   # the second submission's admission fails
    s = ctx.server(MAXX, MAXZ)
    monkeypatch.delenv("CQ_FAULT_ADMIT")
    ids = [s.submit(j) for j in js]
    assert ids[1] == 2
    while any(s.state(i) in (engine.QUEUED, engine.RUNNING) for i in ids):
        s.advance(40)
    assert s.state(ids[1]) == engine.FAILED
    with pytest.raises(engine.CqError) as e:
        s.result(ids[1])
    assert code_of(e.value) == ENOMEM
    for k in (0, 2):
        assert s.state(ids[k]) == engine.DONE
        assert np.array_equal(s.result(ids[k]), alone[k]), k
    # the server keeps serving after the failure
    i = s.submit(js[1])
    while s.state(i) == engine.RUNNING or s.state(i) == engine.QUEUED:
        s.advance(40)
    out = s.result(i)
    assert all(np.array_equal(x, y) for x, y in zip(out, alone[1]))
    s.close()
    assert "CQ_FAULT_ADMIT" not in os.environ
