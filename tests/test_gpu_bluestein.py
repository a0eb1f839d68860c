"""Windows whose padded lengths have a prime factor above 17 (the FFT
stages' largest radix), which the engine transforms by Bluestein's chirp-z
algorithm in LDS, against the FP64 oracle through the C ABI.  cuFFT plans any
length (sgpstd-qmodeling/acps4.cu:871-872); the reference's own default run
has nxx = 321 = 3 107 (input/ac2drealmodeling.txt with geometry/parameters.txt,
window rule acousticsgps4.c:332-336), the batched workload's interior shots
1281 = 3 7 61 and the Q-RTM workload's 2081 (a prime).

Tolerances as for every modeling parity test: gather rel-L2 <= 1e-4 vs FP64;
RTM image <= 1e-3, illumination <= 1e-4."""
import numpy as np
import pytest

import orc
from cqpy import engine, host, synth
from test_gpu_rtm import rtm_job

pytestmark = pytest.mark.gpu

TOL = 1e-4


def largest_prime(n):
    p, m, big = 2, n, 1
    while m > 1:
        while m % p == 0:
            m //= p
            big = p
        p += 1
    return big


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


def run(ctx, jobs):
    b = ctx.batch(jobs)
    b.run()
    out = [b.rec(i) for i in range(len(jobs))]
    b.close()
    return out


SMALL = [
    (43, 28, 5),     # 53 (prime) x 38 (2 19)
    (60, 41, 6),     # 73 x 53, both prime
    (88, 29, 7),     # 103 x 43, both prime
    (301, 160, 10),  # 321 (3 107) x 180: the reference's default window
    (398, 160, 10),  # 419 (prime) x 180
]


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("nxt,nzt,pml", SMALL)
def test_bluestein_windows_vs_oracle(ctx, kind, nxt, nzt, pml):
    nt = 120 if nxt < 200 else 80
    job = synth.small_job(nxt=nxt, nzt=nzt, pml=pml, nt=nt, kind=kind, seed=51)
    print("case kind", kind, "window", job["nxx"], job["nzz"], "nt", nt, flush=True)
    assert max(largest_prime(job["nxx"]), largest_prime(job["nzz"])) > 17
    ref = orc.forward(job)
    got = run(ctx, [job])[0]
    e = engine.rel_l2(got, ref)
    print("rel-L2 vs FP64", e)
    assert e <= TOL


def test_bluestein_mixed_with_smooth_in_one_batch(ctx):
    """A Bluestein window and 7-smooth windows of both kinds in one batch:
    each result equals the shot run alone, bit for bit."""
    jobs = [synth.small_job(nxt=43, nzt=28, pml=5, nt=90, kind=0, seed=52),
            synth.small_job(nxt=40, nzt=24, pml=5, nt=90, kind=1, seed=53),
            synth.small_job(nxt=60, nzt=41, pml=6, nt=90, kind=1, seed=54),
            synth.small_job(nxt=50, nzt=30, pml=6, nt=90, kind=0, seed=55)]
    together = run(ctx, jobs)
    for k, j in enumerate(jobs):
        assert np.array_equal(together[k], run(ctx, [j])[0]), k


@pytest.mark.parametrize("kind", [2, 3])
def test_bluestein_rtm_vs_oracle(ctx, kind):
    job = rtm_job(kind, nxt=43, nzt=28, pml=5, nt=80, snaps=6, seed=56)
    print("case kind", kind, "window", job["nxx"], job["nzz"], flush=True)
    trace, _ = host.revolve_trace(80, 6)
    ref_img, ref_ssg = orc.rtm(job, trace)
    b = ctx.batch([job])
    b.run()
    img, ssg = b.image(0)
    b.close()
    assert engine.rel_l2(ssg, ref_ssg) <= 1e-4
    assert engine.rel_l2(img, ref_img) <= 1e-3


@pytest.mark.parametrize("kind", [0, 1])
def test_natural_batch_window_1281(ctx, kind):
    """An interior shot of the batched workload on its natural window
    (1281 x 680, not widened) against the oracle on the identical grid."""
    job = synth.batch128_jobs(nt=40, nshots=128, kinds=[kind] * 128, shots=[64], widen=False,
                              want_k2=True)[0]
    print("case kind", kind, "window", job["nxx"], job["nzz"], flush=True)
    assert job["nxx"] == 1281 and job["nzz"] == 680
    ref = orc.forward(job)
    got = run(ctx, [job])[0]
    e = engine.rel_l2(got, ref)
    print("rel-L2 vs FP64", e)
    assert e <= TOL


def test_natural_qrtm_window_2081(ctx):
    """The Q-RTM workload's natural window (2081 x 880, 2081 prime): modeling
    of both kinds over a short time axis against the oracle, which runs the
    same lengths by its own Bluestein transform."""
    for kind in (0, 1):
        job = synth.qrtm_jobs(nshots=1, nt=8, snaps=2, kind=kind, widen=False)[0]
        job.update(host.k2d(job["dx"], job["dz"], job["nxx"], job["nzz"], q=True))
        print("case kind", kind, "window", job["nxx"], job["nzz"], flush=True)
        assert job["nxx"] == 2081
        ref = orc.forward(job)
        got = run(ctx, [job])[0]
        e = engine.rel_l2(got, ref)
        print("rel-L2 vs FP64", e, flush=True)
        assert e <= TOL
