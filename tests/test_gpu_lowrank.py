"""The rank-r low-rank form of the viscoacoustic operator on the MI355X
(k_lr_r1 / k_lr_c2 in place of k_q_r1 / k_q_c2), through the C ABI, against
the FP64 oracle running the reference's low-rank apply
(sgpstd-qmodeling/acps4.cu:1337-1459) on the same decomposition.

* The Taylor operator's exact rank-4 decomposition gives the Taylor path's
  gather (the low-rank engine reduces to the live one).
* The fractional-Laplacian operator at rank 4 (the viscoacoustic workload:
  459 x 220, nt = 3000, Q 20-200 plus a Q = 15 anomaly) matches the oracle,
  beside its attenuation-free (acoustic) gather.
* Q-RTM with the low-rank operator; low-rank, Taylor and acoustic shots in
  one batch each equal the shot run alone.
Tolerances: gathers rel-L2 <= 1e-4 vs FP64, images <= 1e-3."""
import numpy as np
import pytest

import orc
from cqpy import engine, host, synth
from test_gpu_rtm import rtm_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


def run(ctx, jobs):
    b = ctx.batch(jobs)
    b.run()
    out = [b.rec(i) if int(j["kind"]) < 2 else b.image(i) for i, j in enumerate(jobs)]
    b.close()
    return out


def test_taylor_decomposition_reduces_to_taylor_path(ctx):
    job = synth.small_job(nxt=40, nzt=24, pml=5, nt=150, kind=1, seed=81)
    lrjob, lr = synth.with_lowrank(job, 4, symbol=host.LR_TAYLOR4, seed=3)
    got, taylor = run(ctx, [lrjob, job])
    ref = orc.forward(job)
    e_t, e_o = engine.rel_l2(got, taylor), engine.rel_l2(got, ref)
    print("low-rank (Taylor operator) vs GPU Taylor path", e_t, "vs FP64 oracle", e_o)
    assert e_t <= 1e-5 and e_o <= 1e-4


@pytest.mark.parametrize("rank", [1, 2, 4, 6, 8])
def test_fraclap_lowrank_vs_oracle_small(ctx, rank):
    job = synth.small_job(nxt=60, nzt=35, pml=8, nt=150, kind=1, seed=82)
    lrjob, lr = synth.with_lowrank(job, rank, seed=5)
    got = run(ctx, [lrjob])[0]
    ref = orc.forward(job, lowrank=lr)
    e = engine.rel_l2(got, ref)
    print("rank", rank, "rel-L2 vs FP64 oracle", e)
    assert e <= 1e-4


def test_viscoacoustic_workload_rank4_full(ctx):
    """459 x 220 padded, nt = 3000, dt = 0.5 ms, rank 4 (FP64 decomposition
    on the host), and the attenuation-free gather of the same shot."""
    qjob = synth.config1_job(nt=3000, kind=1)
    ajob = synth.config1_job(nt=3000, kind=0)
    lrjob, lr = synth.with_lowrank(qjob, 4, seed=11)
    cells = np.random.default_rng(1).choice(qjob["nxx"] * qjob["nzz"], 200, replace=False)
    print("decomposition error (checka_error)", host.lowrank_error(qjob, *lr, cells), flush=True)
    got_q, got_a = run(ctx, [lrjob, ajob])
    ref_q = orc.forward(qjob, lowrank=lr)
    e = engine.rel_l2(got_q, ref_q)
    print("viscoacoustic rank-4 gather rel-L2 vs FP64", e, flush=True)
    assert e <= 1e-4
    assert np.isfinite(got_a).all()
    # attenuation: the viscoacoustic gather carries less energy late in time
    late = slice(1500, 3000)
    assert np.linalg.norm(got_q[late]) < np.linalg.norm(got_a[late])


def test_lowrank_qrtm_vs_oracle(ctx):
    job = rtm_job(3, nt=100, snaps=7, seed=83)
    lrjob, lr = synth.with_lowrank(job, 4, seed=7)
    img, ssg = run(ctx, [lrjob])[0]
    ref_img, ref_ssg = orc.rtm(job, host.revolve_trace(100, 7)[0], lowrank=lr)
    ei, es = engine.rel_l2(img, ref_img), engine.rel_l2(ssg, ref_ssg)
    print("Q-RTM low-rank image", ei, "illumination", es)
    assert es <= 1e-4 and ei <= 1e-3


def test_lowrank_in_mixed_batch_equals_alone(ctx):
    q = synth.small_job(nxt=41, nzt=25, pml=7, nt=90, kind=1, seed=84)
    jobs = [synth.with_lowrank(q, 4, seed=9)[0], q,
            synth.small_job(nxt=40, nzt=24, pml=5, nt=90, kind=0, seed=85),
            synth.with_lowrank(synth.small_job(nxt=43, nzt=28, pml=5, nt=70, kind=1, seed=86), 3, seed=2)[0]]
    together = run(ctx, jobs)
    for k, j in enumerate(jobs):
        assert np.array_equal(together[k], run(ctx, [j])[0]), k
