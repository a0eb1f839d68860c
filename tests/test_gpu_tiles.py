"""Tile shapes, FFT plan variants and the autotuner through the C ABI
(cq_batch_create_tiled, cq_batch_get_tiles, cq_autotune).

The tiles only regroup the same per-line transforms and per-cell updates
among workgroups, and the plan variant only changes the factorisation of
each transform, so every configuration is held to the parity bar against the
FP64 oracle (gathers rel-L2 <= 1e-4, RTM images <= 1e-3, illumination
<= 1e-4): the smallest and largest tiles of every pass, kz-row tiles that do
not divide each other (tk = 4, tk2 = 3), radices of at most 8, receivers and
RTM data injection regrouped by the tiles, and the benchmark's 1296 x 680
window with the radix-8 plan in place of its compiled 16.9.9 / 8.5.17 plans.
The autotuner's choice runs to the same bar and is answered from its
on-disk cache on the second call."""
import numpy as np
import pytest

import orc
from cqpy import engine, host, synth
from test_gpu_rtm import rtm_job

pytestmark = pytest.mark.gpu

TOL = 1e-4

CONFIGS = [
    {"tz": 4, "cw": 2, "cwq": 2, "tk": 1, "tk2": 1},
    {"tz": 16, "cw": 32, "cwq": 32, "tk": 8, "tk2": 16},
    {"tz": 12, "cw": 24, "cwq": 24, "tk": 2, "tk2": 6, "plan": 1},
    {"tz": 8, "cw": 12, "cwq": 16, "tk": 4, "tk2": 3},
]


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def mix():
    jobs = [synth.small_job(nxt=50, nzt=30, pml=6, nt=80, kind=0, seed=31),
            synth.small_job(nxt=43, nzt=37, pml=5, nt=70, kind=1, seed=32),
            rtm_job(2, nt=60, snaps=4, seed=33),
            rtm_job(3, nt=60, snaps=4, seed=34)]
    refs = []
    for j in jobs:
        if j["kind"] < 2:
            refs.append(orc.forward(j))
        else:
            refs.append(orc.rtm(j, host.revolve_trace(int(j["nt"]), int(j["snaps"]))[0]))
    return jobs, refs


def check(b, jobs, refs, label):
    for i, (j, ref) in enumerate(zip(jobs, refs)):
        if j["kind"] < 2:
            e = engine.rel_l2(b.rec(i), ref)
            print(label, "job", i, "gather rel-L2", e)
            assert e <= TOL, (label, i, e)
        else:
            img, ssg = b.image(i)
            ei, es = engine.rel_l2(img, ref[0]), engine.rel_l2(ssg, ref[1])
            print(label, "job", i, "image", ei, "illumination", es)
            assert ei <= 1e-3 and es <= 1e-4, (label, i, ei, es)


@pytest.mark.parametrize("cfg", range(len(CONFIGS)))
def test_tiles_vs_oracle(ctx, mix, cfg):
    jobs, refs = mix
    t = CONFIGS[cfg]
    print("case tiles", t, flush=True)
    b = ctx.batch(jobs, tiles=t)
    got = b.tiles()
    for k, v in t.items():
        assert got[k] == v, (k, got)
    b.run()
    check(b, jobs, refs, str(t))
    b.close()


def test_engine_choice_reported(ctx, mix):
    jobs, _ = mix
    b = ctx.batch(jobs)
    t = b.tiles()
    b.close()
    assert t["tz"] in (4, 8, 12, 16) and t["cw"] in (2, 4, 8, 12, 16, 24, 32)
    assert t["plan"] == 0
    # a partial override keeps the engine's other choices
    b = ctx.batch(jobs, tiles={"cwq": 8})
    u = b.tiles()
    b.close()
    assert u["cwq"] == 8 and u["tz"] == t["tz"] and u["tk"] == t["tk"]


def test_bad_tiles_rejected(ctx, mix):
    jobs, _ = mix
    with pytest.raises(engine.CqError, match="outside its set"):
        ctx.batch(jobs, tiles={"tz": 5})
    with pytest.raises(engine.CqError, match="outside its set"):
        ctx.batch(jobs, tiles={"plan": 2})
    # 16 lines of a 2110-point column exceed the 160 KB workgroup limit
    tall = synth.small_job(nxt=20, nzt=2100, pml=5, nt=5, kind=0, seed=35)
    with pytest.raises(engine.CqError, match="does not fit"):
        ctx.batch([tall], tiles={"cw": 32})


@pytest.mark.parametrize("kind", [0, 1])
def test_benchmark_window_radix8_plan(ctx, kind):
    job = synth.batch128_jobs(nt=40, nshots=128, kinds=[kind] * 128, want_k2=True, shots=[64])[0]
    assert (job["nxx"], job["nzz"]) == (1296, 680)
    print("case benchmark window plan 1 kind", kind, flush=True)
    b = ctx.batch([job], tiles={"plan": 1})
    assert b.tiles()["plan"] == 1
    b.run()
    got = b.rec(0)
    b.close()
    orc.set_threads(16)
    e = engine.rel_l2(got, orc.forward(job))
    print("rel-L2 gather", e)
    assert e <= TOL


@pytest.mark.parametrize("kind", [0, 1])
def test_compressed_twiddles_short_columns(ctx, kind):
    """1080-point columns (8 15 9) in 16-column tiles: eight lines and the
    full 1080-entry twiddle table would take 86.4 KB of LDS, so the plan keeps
    the last stage's first powers only (239 entries, 79.7 KB), as it
    otherwise does only above 2048 points.  Against the FP64 oracle."""
    job = synth.small_job(nxt=40, nzt=1000, pml=40, nt=40, kind=kind, seed=77)
    assert job["nzz"] == 1080
    print("case window", job["nxx"], job["nzz"], "kind", kind, flush=True)
    t = {"cw": 16, "cwq": 16}
    b = ctx.batch([job], tiles=t)
    assert all(b.tiles()[k] == v for k, v in t.items())
    b.run()
    got = b.rec(0)
    b.close()
    e = engine.rel_l2(got, orc.forward(job))
    print("rel-L2 gather", e)
    assert e <= TOL


def test_autotune_and_cache(ctx, mix, tmp_path):
    jobs, refs = mix
    cache = tmp_path / "tiles.txt"
    print("case autotune", flush=True)
    t1, ms1, hit1 = ctx.autotune(jobs, cache)
    print("tuned", t1, ms1, hit1)
    assert not hit1 and ms1 > 0
    lines = [l for l in cache.read_text().splitlines() if not l.startswith("#")]
    assert len(lines) == 1 and lines[0].split("\t")[1].split()[:6] == [str(t1[k]) for k in
                                                                       ("tz", "cw", "cwq", "tk", "tk2", "plan")]
    t2, ms2, hit2 = ctx.autotune(jobs, cache)
    assert hit2 and t2 == t1 and abs(ms2 - ms1) < 1e-5
    # another window set misses the cache and appends its own line
    t3, _, hit3 = ctx.autotune(jobs[:2], cache)
    assert not hit3
    assert len([l for l in cache.read_text().splitlines() if not l.startswith("#")]) == 2
    # the tuned tiles run the jobs to the parity bar, and the jobs passed to
    # the tuner were not advanced by it
    b = ctx.batch(jobs, tiles=t1)
    b.run()
    check(b, jobs, refs, "tuned")
    b.close()
