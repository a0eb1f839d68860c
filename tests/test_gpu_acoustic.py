"""Acoustic modeling on the MI355X through the C ABI, against the FP64 oracle
(the reference's acps4.cu per-step kernels restated on the CPU).

Tolerance: rel-L2 of the gather <= 1e-4 vs FP64 (the modeling bar);
receiver and source indexing are integer and checked bit-exactly."""
import numpy as np
import pytest

import orc
from cqpy import engine, synth

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


SHAPES = [
    # nxt, nzt, pml, nt  -> padded (nxx, nzz) parities
    (40, 24, 5, 200),     # 51 x 34  (3*17 odd, 2*17 even)
    (41, 25, 7, 200),     # 55 x 39  (5*11, 3*13: odd, odd)
    (52, 20, 6, 150),     # 65 x 32  (5*13, 2^5)
    (60, 35, 8, 150),     # 77 x 51  (7*11, 3*17)
    (56, 40, 12, 150),    # 81 x 64  (9*9, 16*4)
]


@pytest.mark.parametrize("nxt,nzt,pml,nt", SHAPES)
def test_small_vs_oracle(ctx, nxt, nzt, pml, nt):
    job = synth.small_job(nxt=nxt, nzt=nzt, pml=pml, nt=nt)
    print("case", nxt, nzt, pml, nt, job["nxx"], job["nzz"], flush=True)
    ref, st = orc.forward(job, want_state=True)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    e = engine.rel_l2(got, ref)
    print("rel-L2 gather", e)
    assert e <= TOL
    ep = engine.rel_l2(b.p(0), st["p"])
    print("rel-L2 final p", ep)
    assert ep <= TOL
    # the gather is p at the receiver cells: last row equals the final field there
    p = b.p(0).reshape(job["nxx"], job["nzz"])
    assert np.array_equal(got[-1], p[job["gxg"], job["gzg"]])
    b.close()


def test_plain_equals_batch(ctx):
    job = synth.small_job(nxt=40, nzt=24, pml=5, nt=80)
    b = ctx.batch([job])
    b.run()
    a = b.rec(0)
    b.close()
    p = ctx.plain_forward(job)
    assert np.array_equal(a, p)


def test_alone_equals_mixed_batch(ctx):
    jobs = [synth.small_job(nxt=40, nzt=24, pml=5, nt=90, seed=1),
            synth.small_job(nxt=61, nzt=32, pml=8, nt=60, seed=2, src_col=10),
            synth.small_job(nxt=45, nzt=40, pml=5, nt=120, seed=3, nrec=17)]
    alone = []
    for j in jobs:
        b = ctx.batch([j])
        b.run()
        alone.append(b.rec(0))
        b.close()
    b = ctx.batch(jobs)
    b.run()
    for i, j in enumerate(jobs):
        assert np.array_equal(b.rec(i), alone[i]), i
    b.close()


def test_duplicate_and_clamped_receivers(ctx):
    job = synth.small_job(nxt=40, nzt=24, pml=5, nt=50)
    job["gxg"] = np.array([5, 5, 20, 45, 45, 10], np.int32)   # duplicates and PML-edge cells
    job["gzg"] = np.array([7, 7, 7, 7, 8, 28], np.int32)
    ref = orc.forward(job)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    assert engine.rel_l2(got, ref) <= TOL
    assert np.array_equal(got[:, 0], got[:, 1])
    b.close()


def test_config1_full_vs_oracle(ctx):
    """Plain acoustic workload at full size: 459 x 220 padded, nt = 3000."""
    job = synth.config1_job(nt=3000)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    b.close()
    orc.set_threads(16)
    ref = orc.forward(job)
    e = engine.rel_l2(got, ref)
    print("config1 rel-L2 vs FP64 oracle:", e)
    assert e <= TOL


def test_bad_inputs_raise(ctx):
    job = synth.small_job(nxt=40, nzt=24, pml=5, nt=10)
    bad = dict(job)
    bad["xsg"] = job["nxx"] + 3
    with pytest.raises(engine.CqError):
        ctx.batch([bad])
    bad = dict(job)
    bad["pml"] = 30          # CPML strips wider than the window
    with pytest.raises(engine.CqError):
        ctx.batch([bad])
    bad = dict(job)
    bad["nt"] = 0
    with pytest.raises(engine.CqError):
        ctx.batch([bad])


@pytest.mark.parametrize("kind", [0, 1])
def test_no_receivers(ctx, kind):
    """A shot with no receivers (nr = 0, an empty gather): it runs alone and
    beside shots that record, its gather is empty, and its final pressure
    field equals the same shot's with receivers bit for bit and the FP64
    oracle's within the modeling tolerance."""
    job = synth.small_job(nxt=40, nzt=24, pml=5, nt=60, kind=kind, seed=71)
    bare = dict(job)
    bare["gxg"] = np.zeros(0, np.int32)
    bare["gzg"] = np.zeros(0, np.int32)
    other = synth.small_job(nxt=41, nzt=25, pml=6, nt=50, kind=1 - kind, seed=72)
    _, st = orc.forward(bare, want_state=True)
    b = ctx.batch([bare, other, job])
    b.run()
    assert b.rec(0).shape == (60, 0)
    p0 = b.p(0)
    assert np.array_equal(p0, b.p(2))
    e = engine.rel_l2(p0, st["p"])
    print("kind", kind, "rel-L2 final p", e)
    assert e <= TOL
    b.close()
    s = ctx.server(64, 48)
    i = s.submit(bare)
    while s.state(i) != engine.DONE:
        s.advance(32)
    assert s.result(i).size == 0
    s.close()


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("nxt,nzt,corner", [(4, 4, False), (5, 6, False), (7, 9, True), (13, 12, True)])
def test_smallest_windows_vs_oracle(ctx, kind, nxt, nzt, corner):
    """The smallest windows the engine accepts (8 x 8 padded with a 2-point
    CPML) up to 17 x 16: single-stage transforms of 8, 9, 11, 13, 16 and 17
    points, tiles wider and deeper than the window; with `corner` the source
    sits in the CPML corner cell (0, 0)."""
    job = synth.small_job(nxt=nxt, nzt=nzt, pml=2, nt=30, kind=kind, seed=76)
    if corner:
        job["xsg"], job["zsg"] = 0, 0
    print("case kind", kind, "window", job["nxx"], job["nzz"], "corner", corner, flush=True)
    ref, st = orc.forward(job, want_state=True)
    b = ctx.batch([job])
    b.run()
    e = engine.rel_l2(b.rec(0), ref)
    ep = engine.rel_l2(b.p(0), st["p"])
    b.close()
    print("rel-L2 gather", e, "final p", ep)
    assert e <= TOL and ep <= TOL


@pytest.mark.parametrize("kind", [0, 1])
def test_widest_window_vs_oracle(ctx, kind):
    """A 5103-point-wide window (27 27 7): two lines of a 4-depth row tile
    and the compressed twiddle table take 93 KB, more than the 80 KB tile
    budget, so the row passes launch with up to 160 KB of LDS (one workgroup
    per CU).  Against the FP64 oracle."""
    job = synth.small_job(nxt=5086, nzt=32, pml=8, nt=20, kind=kind, seed=78)
    assert job["nxx"] == 5103
    print("case window", job["nxx"], job["nzz"], "kind", kind, flush=True)
    ref = orc.forward(job)
    b = ctx.batch([job])
    print("tiles", b.tiles(), flush=True)
    b.run()
    e = engine.rel_l2(b.rec(0), ref)
    b.close()
    print("rel-L2 gather", e)
    assert e <= TOL
