"""cq_forward_static, the allocation-free entry point, through the C ABI:
its gathers equal cq_plain_forward's bit for bit (same passes, same tile
shapes), on two calls in a row whose second shot is smaller (the static
scratch keeps the first call's contents, which the second must not see), for
both modeling kinds; a shot too large for the scratch is refused with an
error, and the entry point still works afterwards."""
import numpy as np
import pytest

from cqpy import engine, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("kind", [0, 1])
def test_static_equals_plain_two_calls(ctx, kind):
    big = synth.small_job(nxt=60, nzt=35, pml=8, nt=160, kind=kind, seed=41)    # 76 x 51
    small = synth.small_job(nxt=40, nzt=24, pml=5, nt=120, kind=kind, seed=42)  # 50 x 34
    for job in (big, small):
        print("kind", kind, "window", job["nxx"], job["nzz"], "nt", job["nt"], flush=True)
        got = ctx.forward_static(job)
        ref = ctx.plain_forward(job)
        assert np.isfinite(got).all()
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref)


def test_static_refuses_oversized_then_recovers(ctx):
    # a 1296 x 680 viscoacoustic window at nt = 5000 needs far more than 64 MiB
    job = synth.batch128_jobs(nt=5000, nshots=128, kinds=[1] * 128, shots=[64])[0]
    with pytest.raises(engine.CqError) as e:
        ctx.forward_static(job)
    assert "static scratch" in str(e.value)
    small = synth.small_job(nxt=40, nzt=24, pml=5, nt=80, kind=0, seed=43)
    assert np.array_equal(ctx.forward_static(small), ctx.plain_forward(small))


def test_static_rejects_rtm(ctx):
    from test_gpu_rtm import rtm_job
    with pytest.raises(engine.CqError):
        ctx.forward_static(rtm_job(2, nt=20, snaps=4, seed=44))
