"""Transforms through composite radices (15, 21, 25, 27) on the MI355X: shots
whose padded windows have row and column lengths that the engine factorises
with a composite stage (the batched workload's edge widths 750, 945, 1050,
1250; depths 60, 84, 100, 300), both modeling kinds, against the FP64 oracle
(the reference's per-step kernels restated on the CPU with its own FFT).

Tolerance: rel-L2 of the gather and of the final pressure <= 1e-4 vs FP64."""
import ctypes as C

import numpy as np
import pytest

import orc
from cqpy import engine, synth

pytestmark = pytest.mark.gpu

TOL = 1e-4
PML = 10


@pytest.fixture(scope="module")
def ctx():
    c = engine.Context(0)
    yield c
    c.close()


def window_job(nxx, nzz, kind, nt=60):
    """A shot on an explicit window of nxx x nzz padded cells (the
    reference's window rule of small_job gives odd widths only)."""
    nx, nz = nxx - 2 * PML, nzz - 2 * PML
    vp, rho, q = synth.layered_model(nx, nz, 10.0, seed=11, nlayers=3, qrange=(20.0, 200.0))
    return synth.make_window_job(kind, vp, rho, nx, nz, 10.0, 10.0, 0.001, nt, 20.0, 0, nx, nx // 2, 2,
                                 np.arange(nx), 2, PML, qt=q, fo=40.0, want_k2=True)


def has_composite(lib, n):
    rad = (C.c_int * 16)()
    ns = lib.cq_fft_plan(n, n, 0, rad, 16, None)
    return ns > 0 and any(r in (15, 21, 25, 27) for r in rad[:ns])


@pytest.mark.parametrize("nxx,nzz", [(750, 60), (1050, 84), (945, 100), (1250, 300)])
@pytest.mark.parametrize("kind", [0, 1])
def test_composite_windows_vs_oracle(ctx, nxx, nzz, kind):
    lib = C.CDLL(engine.lib_path())
    assert has_composite(lib, nxx) and has_composite(lib, nzz)
    job = window_job(nxx, nzz, kind)
    assert (job["nxx"], job["nzz"]) == (nxx, nzz)
    print("case kind", kind, "window", nxx, nzz, flush=True)
    ref, st = orc.forward(job, want_state=True)
    b = ctx.batch([job])
    b.run()
    e = engine.rel_l2(b.rec(0), ref)
    ep = engine.rel_l2(b.p(0), st["p"])
    b.close()
    print("rel-L2 gather %.3e, final p %.3e" % (e, ep))
    assert e <= TOL and ep <= TOL
