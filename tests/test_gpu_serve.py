"""Serving through the C ABI (cq_server_*): jobs submitted while others run
join the same launches (continuous batching).  Each result equals the job
run alone on a server of the same window limits (bit-exact: same tiles and
plans), modeling gathers match the FP64 oracle (rel-L2 <= 1e-4), and the
admission limits hold."""
import numpy as np
import pytest

import orc
from cqpy import engine, synth
from test_gpu_rtm import rtm_job

pytestmark = pytest.mark.gpu

MAXX, MAXZ = 72, 48
TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


def job_mix():
    return [synth.small_job(nxt=40, nzt=24, pml=5, nt=120, kind=0, seed=21),
            synth.small_job(nxt=41, nzt=25, pml=7, nt=100, kind=1, seed=22),
            rtm_job(3, nt=80, snaps=5, seed=23),
            synth.small_job(nxt=52, nzt=20, pml=6, nt=90, kind=0, seed=24),
            rtm_job(2, nt=60, snaps=60, seed=25)]


def same(a, b):
    if isinstance(a, tuple):
        return all(np.array_equal(x, y) for x, y in zip(a, b))
    return np.array_equal(a, b)


def run_alone(ctx, job, mx=MAXX, mz=MAXZ):
    s = ctx.server(mx, mz)
    i = s.submit(job)
    while s.state(i) != engine.DONE:
        s.advance(64)
    out = s.result(i)
    s.close()
    return out


def test_serve_staggered_equals_alone(ctx):
    js = job_mix()
    alone = [run_alone(ctx, j) for j in js]
    s = ctx.server(MAXX, MAXZ)
    ids = [s.submit(js[0])]
    s.advance(30)
    ids += [s.submit(js[1]), s.submit(js[2])]
    s.advance(17)
    ids += [s.submit(js[3]), s.submit(js[4])]
    while any(s.state(i) != engine.DONE for i in ids):
        s.advance(40)
    for k, i in enumerate(ids):
        print("job", k, "steps (submitted, started, finished)", s.steps(i), flush=True)
        assert same(s.result(i), alone[k]), k
    st = [s.steps(i) for i in ids]
    # later jobs started while earlier ones were still running
    assert st[1][1] < st[0][2] and st[3][1] < st[2][2]
    for k in (0, 1, 3):
        e = engine.rel_l2(s.result(ids[k]), orc.forward(js[k]))
        print("job", k, "rel-L2 vs FP64 oracle", e)
        assert e <= TOL
    for i in ids:
        s.release(i)
    s.close()


def test_serve_job_limit_runs_one_at_a_time(ctx):
    js = job_mix()[:3]
    s = ctx.server(MAXX, MAXZ, max_jobs=1)
    ids = [s.submit(j) for j in js]
    running, queued = s.advance(1)
    assert running == 1 and queued == 2
    while any(s.state(i) != engine.DONE for i in ids):
        running, queued = s.advance(25)
        assert running <= 1
    st = [s.steps(i) for i in ids]
    assert st[1][1] >= st[0][2] and st[2][1] >= st[1][2]
    for k, i in enumerate(ids):
        assert same(s.result(i), run_alone(ctx, js[k])), k
    s.close()


def test_serve_window_with_tiles_of_its_own(ctx):
    """Windows 409 deep (a prime: Bluestein, 840-point inner transforms) on a
    server for windows up to 64 x 600, whose 24-column tiles would need 97 KB
    of LDS for them (12 inner lines): they run with tiles of their own,
    launched beside the server's, next to jobs on the server's tiles.  Every
    result equals the job alone on a server of the same limits, and the
    gathers match the FP64 oracle."""
    mx, mz = 64, 600
    jobs = [synth.small_job(nxt=40, nzt=399, pml=5, nt=40, kind=k, seed=61 + k) for k in (0, 1)]
    jobs += [synth.small_job(nxt=40, nzt=24, pml=5, nt=60, kind=0, seed=63),
             synth.small_job(nxt=41, nzt=25, pml=7, nt=50, kind=1, seed=64)]
    for j in jobs:
        print("case window", j["nxx"], j["nzz"], "kind", j["kind"], flush=True)
    alone = [run_alone(ctx, j, mx, mz) for j in jobs]
    s = ctx.server(mx, mz)
    ids = [s.submit(j) for j in jobs]
    while any(s.state(i) != engine.DONE for i in ids):
        s.advance(16)
    for k, i in enumerate(ids):
        assert same(s.result(i), alone[k]), k
    s.close()
    for k in (0, 1):
        e = engine.rel_l2(alone[k], orc.forward(jobs[k]))
        print("kind", k, "rel-L2 vs FP64", e)
        assert e <= TOL


def test_serve_rejects_oversized_and_unfinished(ctx):
    s = ctx.server(MAXX, MAXZ)
    big = synth.small_job(nxt=80, nzt=24, pml=5, nt=10, kind=0, seed=26)
    with pytest.raises(engine.CqError):
        s.submit(big)
    i = s.submit(job_mix()[0])
    with pytest.raises(engine.CqError):
        s.result(i)
    with pytest.raises(engine.CqError):
        s.release(i)
    s.close()


def test_serve_run_cells_shortest_first_equals_alone(ctx):
    """run_cells = 1 (one job per global step): jobs run shortest remaining
    work first; a short job submitted later pauses the long Q-RTM job in the
    middle of its revolve schedule, which then resumes where it stopped.
    Every result equals the job run alone, bit for bit."""
    js = job_mix()
    alone = [run_alone(ctx, j) for j in js]
    s = ctx.server(MAXX, MAXZ, run_cells=1)
    q = s.submit(js[2])                  # Q-RTM (revolve checkpoints)
    running, queued = s.advance(40)
    assert running == 1 and queued == 0
    a = s.submit(js[0])                  # shorter: pauses the Q-RTM job
    others = [s.submit(js[k]) for k in (1, 3, 4)]
    ids = [a, q] + others
    while any(s.state(i) != engine.DONE for i in ids):
        s.advance(25)
    st = {i: s.steps(i) for i in ids}
    for i in ids:
        print("job", i, "steps (submitted, started, finished)", st[i], flush=True)
    # the short acoustic job finished before the Q-RTM job it interrupted
    assert st[a][2] < st[q][2]
    # one job at a time: finish steps are distinct
    assert len({st[i][2] for i in ids}) == len(ids)
    for k, i in zip([0, 2, 1, 3, 4], ids):
        assert same(s.result(i), alone[k]), k
    s.close()
