"""Graph-captured steps (cq_batch_use_graphs) through the C ABI: a batch of
acoustic, viscoacoustic, low-rank, RTM and Q-RTM shots of different nt run
with spans of k steps replayed as HIP graphs gives the same gathers,
wavefields and images as direct launches, bit for bit.  The graphs are
replayed many times, with direct steps, host reads and synchronisations
between the replays (the counters must be reloaded each time), spans are cut
at RTM checkpoint copies, and switching graphs on and off mid-run changes
nothing."""
import numpy as np
import pytest

from cqpy import engine, host, synth
from test_gpu_rtm import rtm_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


def jobs():
    lr = synth.with_lowrank(synth.small_job(nxt=44, nzt=26, pml=5, nt=70, kind=1, seed=43), 3, host.LR_FRACLAP, 5)[0]
    return [synth.small_job(nxt=40, nzt=24, pml=5, nt=90, kind=0, seed=41),
            synth.small_job(nxt=41, nzt=25, pml=7, nt=75, kind=1, seed=42),
            lr,
            rtm_job(2, nt=50, snaps=4, seed=44),
            rtm_job(3, nt=40, snaps=3, seed=45)]


def results(b, js):
    out = []
    for i, j in enumerate(js):
        if j["kind"] < 2:
            out.append((b.rec(i), b.p(i)))
        else:
            out.append(b.image(i))
    return out


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_graphs_equal_direct_launches(ctx):
    js = jobs()
    b = ctx.batch(js)
    b.run()
    ref = results(b, js)
    b.close()
    for k in (1, 4, 16):
        print("case graph span", k, flush=True)
        b = ctx.batch(js)
        b.use_graphs(k)
        b.run()
        got = results(b, js)
        b.close()
        assert all(same(x, y) for x, y in zip(ref, got)), k


def test_graph_replays_interleaved_with_direct_steps(ctx):
    js = jobs()
    b = ctx.batch(js)
    b.run()
    ref = results(b, js)
    b.close()
    print("case graph replays with direct steps between", flush=True)
    b = ctx.batch(js)
    b.use_graphs(8)
    rng = np.random.default_rng(3)
    mid = []
    # the batch's longest schedule is the RTM shot's 194 steps: stop short of it
    for _ in range(100):
        if b.time() >= 150:
            break
        n = int(rng.integers(1, 30))
        b.step(n)
        b.sync()
        mid.append(b.p(0).copy())          # host read between replays
        if rng.uniform() < 0.3:
            b.use_graphs(0)                # a few direct steps
            b.step(int(rng.integers(1, 5)))
            b.use_graphs(int(rng.choice([2, 8])))
    b.run()
    got = results(b, js)
    b.close()
    assert all(same(x, y) for x, y in zip(ref, got))
    assert all(np.isfinite(m).all() for m in mid)
