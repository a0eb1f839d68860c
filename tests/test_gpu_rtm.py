"""RTM and Q-RTM with revolve checkpointing on the MI355X through the C ABI,
against the FP64 oracle replaying the same revolve schedule
(sgpstd-qrtm/acps4.cu:447-945, sgpstd-qrtm/acsglw.cu:470-1079).

Tolerance: rel-L2 of the image <= 1e-3 and of the illumination <= 1e-4."""
import numpy as np
import pytest

import orc
from cqpy import engine, host, synth

pytestmark = pytest.mark.gpu

TOL_IMG = 1e-3
TOL_SSG = 1e-4


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


def rtm_job(kind, nxt=40, nzt=24, pml=5, nt=120, snaps=6, seed=7, eta=1e-20):
    """Observed data from the FP64 oracle on one model, migration on another
    (different seed), direct wave removed as the reference's driver does."""
    qk = 1 if kind == 3 else 0
    jt = synth.small_job(nxt=nxt, nzt=nzt, pml=pml, nt=nt, kind=qk, seed=seed)
    obs = orc.forward(jt).astype(np.float32)
    jm = synth.small_job(nxt=nxt, nzt=nzt, pml=pml, nt=nt, kind=kind, seed=seed + 100)
    nr = obs.shape[1]
    gxp = (np.linspace(0, nxt - 1, nr) * jm["dx"]).astype(np.float32)
    gzp = np.full(nr, 2 * jm["dx"], np.float32)
    obs = host.removedirect(obs, nt, nr, jm["dt"], jm["fd"], (nxt // 2) * jm["dx"], 2 * jm["dx"],
                            gxp, gzp, 1500.0).reshape(nt, nr)
    jm.update(obs=obs, snaps=snaps, etas=eta, etar=eta)
    return jm


@pytest.mark.parametrize("kind", [2, 3])
@pytest.mark.parametrize("nt,snaps", [(120, 6), (90, 40), (64, 1), (120, 64)])
def test_rtm_vs_oracle(ctx, kind, nt, snaps):
    job = rtm_job(kind, nt=nt, snaps=snaps)
    print("case kind", kind, "nt", nt, "snaps", snaps, job["nxx"], job["nzz"], flush=True)
    trace, info = host.revolve_trace(nt, snaps)
    assert trace[-1][0] == 5, info
    ref_img, ref_ssg = orc.rtm(job, trace)
    b = ctx.batch([job])
    b.run()
    img, ssg = b.image(0)
    b.close()
    ei = engine.rel_l2(img, ref_img)
    es = engine.rel_l2(ssg, ref_ssg)
    print("rel-L2 image", ei, "illumination", es)
    assert es <= TOL_SSG
    assert ei <= TOL_IMG


def test_rtm_alone_equals_batch(ctx):
    jobs = [rtm_job(2, seed=3), rtm_job(2, seed=4, nxt=45, nzt=40)]
    alone = []
    for j in jobs:
        b = ctx.batch([j])
        b.run()
        alone.append(b.image(0))
        b.close()
    b = ctx.batch(jobs)
    b.run()
    for i in range(len(jobs)):
        img, ssg = b.image(i)
        assert np.array_equal(img, alone[i][0]) and np.array_equal(ssg, alone[i][1])
    b.close()


def test_rtm_linear_in_data(ctx):
    """The receiver wavefield, hence the image, is linear in the data."""
    j1 = rtm_job(2, seed=5)
    j2 = dict(j1)
    j2["obs"] = 2.0 * j1["obs"]
    b = ctx.batch([j1, j2])
    b.run()
    i1, _ = b.image(0)
    i2, _ = b.image(1)
    b.close()
    assert engine.rel_l2(i2, 2.0 * i1) <= 1e-6


def test_rtm_mixed_schedules_and_modeling_in_one_batch(ctx):
    """Shots with different nt, checkpoint budgets and kinds (RTM, Q-RTM and a
    forward-modeling shot) advance in the same launches, each on its own
    revolve schedule: every result equals the shot run alone.  Three RTM
    shots put two of them in one launch group (one stream, interleaved
    checkpoint copies)."""
    jobs = [rtm_job(2, nt=60, snaps=4, seed=11), rtm_job(3, nt=90, snaps=7, seed=12),
            rtm_job(2, nt=75, snaps=75 // 2, seed=13), rtm_job(2, nt=50, snaps=3, seed=15)]
    fwd = synth.small_job(nxt=40, nzt=24, pml=5, nt=50, kind=1, seed=14)
    alone = []
    for j in jobs:
        b = ctx.batch([j])
        b.run()
        alone.append(b.image(0))
        b.close()
    b = ctx.batch([fwd])
    b.run()
    rec_alone = b.rec(0)
    b.close()
    b = ctx.batch(jobs + [fwd])
    b.run()
    for i in range(len(jobs)):
        img, ssg = b.image(i)
        assert np.array_equal(img, alone[i][0]) and np.array_equal(ssg, alone[i][1]), i
    assert np.array_equal(b.rec(len(jobs)), rec_alone)
    b.close()


def test_qrtm_long_lines_vs_oracle(ctx):
    """Q-RTM on a 1575-column window (one-at-a-time symbol inverses in the
    kz-row pass, stabilised multiplier on both wavefields)."""
    job = rtm_job(3, nxt=1555, nzt=30, pml=10, nt=40, snaps=8, seed=17)
    assert job["nxx"] == 1575
    print("case long lines Q-RTM", job["nxx"], job["nzz"], flush=True)
    trace, info = host.revolve_trace(40, 8)
    ref_img, ref_ssg = orc.rtm(job, trace)
    b = ctx.batch([job])
    b.run()
    img, ssg = b.image(0)
    b.close()
    ei, es = engine.rel_l2(img, ref_img), engine.rel_l2(ssg, ref_ssg)
    print("rel-L2 image", ei, "illumination", es)
    assert es <= TOL_SSG and ei <= TOL_IMG


def test_mixed_batch_in_uneven_chunks_equals_run(ctx):
    """The launch groups' streams fork and join at every cq_batch_step call:
    advancing a mixed batch in uneven chunks gives the results of one run."""
    jobs = [rtm_job(2, nt=40, snaps=3, seed=21), rtm_job(3, nt=50, snaps=5, seed=22),
            synth.small_job(nxt=40, nzt=24, pml=5, nt=45, kind=0, seed=23),
            synth.small_job(nxt=40, nzt=24, pml=5, nt=35, kind=1, seed=24)]
    b = ctx.batch(jobs)
    b.run()
    ref = [b.image(0), b.image(1), b.rec(2), b.rec(3)]
    b.close()
    b = ctx.batch(jobs)
    for k in [1, 7, 3, 13, 2, 29, 5] * 20:
        b.step(k)
    b.sync()
    got = [b.image(0), b.image(1), b.rec(2), b.rec(3)]
    b.close()
    for i in range(2):
        assert np.array_equal(got[i][0], ref[i][0]) and np.array_equal(got[i][1], ref[i][1]), i
    assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


@pytest.mark.parametrize("kind", [2, 3])
def test_rtm_no_receivers(ctx, kind):
    """An RTM shot with no receivers (nr = 0: no observed data): the receiver
    wavefield stays zero, so the image is zero, and the illumination equals
    the same shot's with receivers bit for bit (it depends on the source
    wavefield alone)."""
    job = rtm_job(kind, nt=60, snaps=5, seed=73)
    bare = dict(job)
    bare["gxg"] = np.zeros(0, np.int32)
    bare["gzg"] = np.zeros(0, np.int32)
    bare["obs"] = np.zeros((60, 0), np.float32)
    b = ctx.batch([bare, job])
    b.run()
    img0, ssg0 = b.image(0)
    img1, ssg1 = b.image(1)
    b.close()
    assert not img0.any()
    assert ssg0.any() and np.array_equal(ssg0, ssg1)
    assert img1.any()


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("nt", [1, 2, 3])
def test_shortest_time_axes_vs_oracle(ctx, kind, nt):
    """Time axes of one to three samples (a one-step revolve schedule has no
    forward sweep: its first turn is the receiver step of sample 0) against
    the FP64 oracle; an all-zero reference must be matched by zeros."""
    if kind < 2:
        job = synth.small_job(nxt=40, nzt=24, pml=5, nt=nt, kind=kind, seed=74)
        ref = orc.forward(job)
    else:
        job = rtm_job(kind, nt=nt, snaps=2, seed=75)
        ref = orc.rtm(job, host.revolve_trace(nt, 2)[0])
    print("case kind", kind, "nt", nt, flush=True)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0) if kind < 2 else b.image(0)
    b.close()
    checks = [(got, ref, 1e-4)] if kind < 2 else [(got[0], ref[0], TOL_IMG), (got[1], ref[1], TOL_SSG)]
    for g, r, tol in checks:
        if not np.any(r):
            assert not np.any(g)
            continue
        e = engine.rel_l2(g, r)
        print("rel-L2", e)
        assert e <= tol
