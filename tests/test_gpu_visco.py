"""Viscoacoustic (fractional-Laplacian constant-Q) modeling on the MI355X
through the C ABI, against the FP64 oracle (the reference's acsglw.cu per-step
kernels restated on the CPU).

Tolerance: rel-L2 of the gather <= 1e-4 vs FP64."""
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
    (40, 24, 5, 200),     # 51 x 34  (odd x, even z)
    (41, 25, 7, 200),     # 55 x 39  (odd, odd)
    (52, 20, 6, 150),     # 65 x 32
    (56, 40, 12, 150),    # 81 x 64
]


@pytest.mark.parametrize("nxt,nzt,pml,nt", SHAPES)
def test_visco_small_vs_oracle(ctx, nxt, nzt, pml, nt):
    job = synth.small_job(nxt=nxt, nzt=nzt, pml=pml, nt=nt, kind=1)
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
    b.close()


def test_visco_attenuates(ctx):
    """Same shot, acoustic vs viscoacoustic: Q < 200 must lower the late-time
    energy of the gather."""
    ja = synth.small_job(nxt=56, nzt=40, pml=12, nt=300, kind=0)
    jq = synth.small_job(nxt=56, nzt=40, pml=12, nt=300, kind=1)
    b = ctx.batch([ja, jq])
    b.run()
    ra, rq = b.rec(0), b.rec(1)
    b.close()
    late = slice(150, 300)
    assert np.linalg.norm(rq[late]) < np.linalg.norm(ra[late])


def test_q_infinity_is_acoustic(ctx):
    """gamma = 1/(pi Q) -> 0 gives y1 = y3 = 1, y2 = y4 = 0, lam = 1, tao -> 0:
    the viscoacoustic step reduces to the acoustic one."""
    ja = synth.small_job(nxt=52, nzt=20, pml=6, nt=150, kind=0)
    jq = synth.small_job(nxt=52, nzt=20, pml=6, nt=150, kind=1)
    n = jq["nxx"] * jq["nzz"]
    jq["gam"] = np.zeros(n, np.float32)
    jq["lam"] = np.ones(n, np.float32)
    jq["tao"] = np.zeros(n, np.float32)
    b = ctx.batch([ja, jq])
    b.run()
    e = engine.rel_l2(b.rec(1), b.rec(0))
    b.close()
    print("Q=inf vs acoustic rel-L2", e)
    assert e <= 1e-5


def test_mixed_kinds_alone_equals_batch(ctx):
    jobs = [synth.small_job(nxt=40, nzt=24, pml=5, nt=90, seed=1, kind=1),
            synth.small_job(nxt=61, nzt=32, pml=8, nt=60, seed=2, src_col=10, kind=0),
            synth.small_job(nxt=45, nzt=40, pml=5, nt=120, seed=3, nrec=17, kind=1)]
    alone = []
    for j in jobs:
        b = ctx.batch([j])
        b.run()
        alone.append(b.rec(0))
        b.close()
    b = ctx.batch(jobs)
    b.run()
    for i in range(len(jobs)):
        assert np.array_equal(b.rec(i), alone[i]), i
    b.close()


def test_config2_full_vs_oracle(ctx):
    """Viscoacoustic workload at full size: 459 x 220 padded, nt = 3000,
    Q 20-200 per layer with a Q=15 anomaly, fo = fd = 15 Hz."""
    job = synth.config1_job(nt=3000, kind=1)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    b.close()
    orc.set_threads(16)
    ref = orc.forward(job)
    e = engine.rel_l2(got, ref)
    print("config2 rel-L2 vs FP64 oracle:", e)
    assert e <= TOL


def test_visco_long_lines_vs_oracle(ctx):
    """A 1575-column window: the kz-row pass no longer fits its three symbol
    inverses in one workgroup's LDS budget and inverts them one at a time."""
    job = synth.small_job(nxt=1555, nzt=40, pml=10, nt=40, kind=1, seed=31)
    assert job["nxx"] == 1575
    print("case long lines", job["nxx"], job["nzz"], flush=True)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    b.close()
    e = engine.rel_l2(got, orc.forward(job))
    print("rel-L2 gather", e)
    assert e <= TOL


@pytest.mark.parametrize("kind", [0, 1])
def test_very_long_lines_vs_oracle(ctx, kind):
    """A 2025-column window: row tiles of 4 depths (two packed lines) fit the
    LDS budget where 8 do not."""
    job = synth.small_job(nxt=2005, nzt=24, pml=10, nt=30, kind=kind, seed=33)
    assert job["nxx"] == 2025
    print("case very long lines kind", kind, job["nxx"], job["nzz"], flush=True)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    b.close()
    e = engine.rel_l2(got, orc.forward(job))
    print("rel-L2 gather", e)
    assert e <= TOL


def test_single_kz_row_tiles_vs_oracle(ctx):
    """A 2125-column window: neither two kz rows with their three symbols
    (6 lines) nor two rows one symbol at a time (4 lines) fit the LDS budget,
    so the kz-row pass takes one kz row per workgroup (3 lines, div and p
    transformed together)."""
    job = synth.small_job(nxt=2104, nzt=24, pml=10, nt=30, kind=1, seed=35)
    assert job["nxx"] == 2125
    print("case single kz rows", job["nxx"], job["nzz"], flush=True)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    b.close()
    e = engine.rel_l2(got, orc.forward(job))
    print("rel-L2 gather", e)
    assert e <= TOL


@pytest.mark.parametrize("kind", [0, 1])
def test_tall_window_vs_oracle(ctx, kind):
    """A 640-depth window: column tiles of 24 columns (12 packed lines; the
    viscoacoustic column pass then runs its partial sums in LDS, two halves
    of 12 columns), and an odd width whose last tile is partial."""
    job = synth.small_job(nxt=44, nzt=620, pml=10, nt=40, kind=kind, seed=37)
    print("case tall window kind", kind, job["nxx"], job["nzz"], flush=True)
    assert job["nzz"] == 640 and job["nxx"] == 65
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    b.close()
    e = engine.rel_l2(got, orc.forward(job))
    print("rel-L2 gather", e)
    assert e <= TOL


def _window(kind, nx, nz, seed):
    """A shot on an explicit nx x nz window (10-point CPML) of a layered
    model, receivers on every fourth surface column."""
    dx, pml = 10.0, 10
    vp, rho, q = synth.layered_model(nx, nz, dx, seed=seed, nlayers=3, qrange=(20.0, 200.0))
    return synth.make_window_job(kind, vp, rho, nx, nz, dx, dx, 0.001, 30, 25.0, 0, nx, nx // 2, 2,
                                 np.arange(0, nx, 4), 2, pml, qt=q, fo=50.0, want_k2=True)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("shape", [(2080, 40), (45, 860)])
def test_qrtm_window_lengths_vs_oracle(ctx, kind, shape):
    """The lengths of the Q-RTM workload's windows, which run FFT stages
    compiled for them: 2100 = 4*3*5*5*7 wide (row and kz-row transforms) and
    880 = 16*5*11 deep (column transforms)."""
    job = _window(kind, shape[0], shape[1], seed=41 + kind)
    print("case Q-RTM lengths kind", kind, job["nxx"], job["nzz"], flush=True)
    assert (job["nxx"], job["nzz"]) in ((2100, 60), (65, 880))
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    b.close()
    e = engine.rel_l2(got, orc.forward(job))
    print("rel-L2 gather", e)
    assert e <= TOL


@pytest.mark.parametrize("kind", [0, 1])
def test_benchmark_window_vs_oracle(ctx, kind):
    """One interior shot of the batched benchmark workload at its full
    1296 x 680 padded window (12-depth row tiles, 24-column tiles, two kz
    rows per kz-row workgroup), 150 steps, against the FP64 oracle."""
    job = synth.batch128_jobs(nt=150, nshots=128, kinds=[kind] * 128, want_k2=True, shots=[64])[0]
    print("case benchmark window kind", kind, job["nxx"], job["nzz"], flush=True)
    assert (job["nxx"], job["nzz"]) == (1296, 680)
    b = ctx.batch([job])
    b.run()
    got = b.rec(0)
    b.close()
    orc.set_threads(16)
    e = engine.rel_l2(got, orc.forward(job))
    print("rel-L2 gather", e)
    assert e <= TOL
