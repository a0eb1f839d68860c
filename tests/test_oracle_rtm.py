"""Pins of the oracle's RTM restatement that hold without the reference:
- the illumination of the first forward sweep at a receiver cell equals the
  sum of squares of the modeled gather there (the acoustic RTM source step is
  the modeling step, sgpstd-qrtm/acps4.cu:1623-1659 vs acps4.cu:1462-1498);
- the image is linear in the observed data;
- the image vanishes for zero data."""
import numpy as np

import orc
from cqpy import host, synth


def job(kind=2, nt=60, snaps=4, seed=7):
    j = synth.small_job(nxt=40, nzt=24, pml=5, nt=nt, kind=kind, seed=seed)
    rng = np.random.default_rng(seed)
    j.update(obs=rng.standard_normal((nt, len(j["gxg"]))).astype(np.float32), snaps=snaps,
             etas=1e-20, etar=1e-20)
    return j


def test_illumination_equals_gather_energy():
    j = job()
    j0 = dict(j)
    j0["kind"] = 0
    rec = orc.forward(j0)
    trace, _ = host.revolve_trace(j["nt"], j["snaps"])
    img, ssg = orc.rtm(j, trace)
    pml = j["pml"]
    for i in range(0, len(j["gxg"]), 7):
        gx, gz = j["gxg"][i], j["gzg"][i]
        if pml <= gx < j["nxx"] - pml and pml <= gz < j["nzz"] - pml:
            assert np.isclose(ssg[gx - pml, gz - pml], np.sum(rec[:, i] ** 2), rtol=1e-12)


def test_image_linear_and_zero():
    j = job()
    trace, _ = host.revolve_trace(j["nt"], j["snaps"])
    i1, s1 = orc.rtm(j, trace)
    j2 = dict(j)
    j2["obs"] = -2.0 * j["obs"]
    i2, s2 = orc.rtm(j2, trace)
    assert np.allclose(i2, -2.0 * i1, rtol=1e-10, atol=1e-12 * np.abs(i1).max())
    assert np.array_equal(s1, s2)
    j3 = dict(j)
    j3["obs"] = np.zeros_like(j["obs"])
    i3, _ = orc.rtm(j3, trace)
    assert np.abs(i3).max() == 0.0
