"""Pins the oracle's time step (oracle/cq_oracle.c: orc_forward, orc_rtm) by
mathematics rather than by restating it.

* Single Fourier modes against closed forms derived by hand from the
  continuous derivative of a cosine and the reference's table rule
  (sgpstd-qmodeling/funca.c:342-427, acps4.cu:508-645): the staggered
  derivative of cos(k x + phi) is -k sin(k (x -/+ d/2) + phi) for even n;
  for odd n the mirror-index quirk gives -k cos(pi/2n) sin(k (x -/+ d/2) + phi
  +/- pi/2n).  One step is then a 2 x 2 recurrence on the (p, v) amplitudes.
* Every mode at once: a homogeneous periodic medium advanced per mode in
  Fourier space (tests/spectral.py: numpy FFTs, Hermitian multipliers in
  closed form, no index-selected complex kernels, no real-part extraction)
  for the acoustic step, CPML memory variables with constant coefficients,
  the viscoacoustic step (acsglw.cu:511-683), RTM and Q-RTM with the
  stabilised multiplier (sgpstd-qrtm/acsglw.cu:1895-1924), data replacement
  with duplicate receivers and the image / illumination sums.
* Invariants: Q -> infinity reduces the viscoacoustic oracle to the acoustic
  one on a heterogeneous model; a constant-Q medium damps a plane wave at the
  rate its Q implies; a revolve-reconstructed RTM equals full storage.

The FP64 oracle and the closed forms differ only by the float32 rounding of
the reference's coefficient tables (about 1e-7 relative per operator), hence
the 1e-5 tolerances on fields (1e-4 on gathers, see TOL_G); every plausible mistake tried on the oracle (a sign, a swapped
shift table, a dropped 1/12 or 1/N, an off-by-one mirror index, a CPML term
in the wrong place, the wrong data row or time index) moves these results by
1e-3 or more.
"""
import numpy as np
import pytest

import orc
import spectral
from cqpy import host, synth

TOL = 1e-5
# Gathers of a point source sample the dispersive tail of the wavefield,
# where the float32 tables' rounding grows to a few 1e-5 relative (with the
# exact table values the residual is 2e-6); a missing odd-n quirk gives 5e-2
# and a 1e-4 change of the velocity 2e-3.
TOL_G = 1e-4


def homog_job(nxx, nzz, dx=10.0, dz=10.0, dt=1e-3, nt=20, vp=1500.0, rho=1024.0, kind=0,
              cpml=None, q=None, src=(5, 7), rik=None, recs=((3, 4),)):
    """A job dict (the oracle's inputs) for a homogeneous periodic window:
    uniform fields, constant CPML coefficients (or none), the reference's
    shift and |k| tables (pinned bit-exact against its host code)."""
    N = nxx * nzz
    f32 = lambda v, n=N: np.full(n, v, np.float32)
    job = dict(kind=kind, nxx=nxx, nzz=nzz, pml=2, nt=nt, dt=dt, dx=dx, dz=dz)
    job["vp"], job["rho"] = f32(vp), f32(rho)
    job["bbx"], job["bbz"] = host.bbx_bbz(nxx, nzz, job["rho"])
    c = cpml or dict(ia=0.0, ib=1.0, ha=0.0, hb=1.0)
    for ax, n in (("x", nxx), ("z", nzz)):
        job["ia" + ax], job["ib" + ax] = f32(c["ia"], n), f32(c["ib"], n)
        job["ha" + ax], job["hb" + ax] = f32(c["ha"], n), f32(c["hb"], n)
    if q is not None:
        for k in ("gam", "lam", "vwd", "tao"):
            job[k] = f32(q[k])
    job.update(host.shift_coeffs(dx, dz, nxx, nzz))
    job.update(host.k2d(dx, dz, nxx, nzz, q=True))
    job["rik"] = np.zeros(nt, np.float32) if rik is None else np.asarray(rik, np.float32)
    job["xsg"], job["zsg"] = src
    job["gxg"] = np.array([r[0] for r in recs], np.int32)
    job["gzg"] = np.array([r[1] for r in recs], np.int32)
    return job


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---------------------------------------------------------------------------
# single modes, closed form
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,m", [(24, 3), (24, 11), (25, 4), (25, 12), (27, 13)])
@pytest.mark.parametrize("axis", ["x", "z"])
def test_single_mode_closed_form(n, m, axis):
    """p0 = cos(2 pi m i / n + phi) along one axis, v0 = 0, no source, no
    CPML: after t steps p = Re(P_t e^{i a}), v = Re(V_t e^{i a}) with
    D- = i k (e^{-i th} + e^{-i th'}) / 2, D+ = i k (e^{i th} + e^{i th'}) / 2,
    th = pi m / n, th' = th (even n) or th - pi / n (odd n), and
      P_{t+1} = P_t + rho v^2 dt (1 - v^2 dt^2 k^2 / 12) D- V_t,
      V_{t+1} = V_t + dt (1/rho) D+ P_{t+1}."""
    other = 16
    nxx, nzz = (n, other) if axis == "x" else (other, n)
    d = 10.0 if axis == "x" else 7.5
    dt, vp, rho, nt = 1.2e-3, 2000.0, 1024.0, 25
    job = homog_job(nxx, nzz, dx=10.0, dz=7.5, dt=dt, nt=nt, vp=vp, rho=rho)
    phi = 0.3
    i = np.arange(n)
    prof = np.cos(2 * np.pi * m * i / n + phi)
    p0 = np.outer(prof, np.ones(nzz)) if axis == "x" else np.outer(np.ones(nxx), prof)
    rec, st = orc.forward(job, want_state=True, init={"p": p0})
    # closed form (k and the phase from the reference's float grid rules)
    k = float(np.float32(m) * np.float32(2 * spectral.PI_REF / (n * d)))
    th = float(np.float32(m * spectral.PI_REF / n))
    thp = th if n % 2 == 0 else float(np.float32((m - 1) * spectral.PI_REF / n))
    dm = 1j * k * (np.exp(-1j * th) + np.exp(-1j * thp)) / 2
    dp = 1j * k * (np.exp(1j * th) + np.exp(1j * thp)) / 2
    P, V = np.exp(1j * phi), 0j
    pmax = vmax = 0.0
    for _ in range(nt):
        P = P + rho * vp * vp * dt * (1 - (vp * dt * k) ** 2 / 12) * dm * V
        V = V + dt / rho * dp * P
        pmax, vmax = max(pmax, abs(P)), max(vmax, abs(V))
    e = np.exp(2j * np.pi * m * i / n)
    p_cf, v_cf = (P * e).real, (V * e).real
    # errors relative to the amplitudes the mode reaches over the run (the
    # final velocity may be near a zero crossing)
    scale = np.linalg.norm(np.cos(2 * np.pi * m * i / n))
    rel_p = lambda a: np.linalg.norm(a - p_cf) / (pmax * scale)
    rel_v = lambda a: np.linalg.norm(a - v_cf) / (vmax * scale)
    p = st["p"].reshape(nxx, nzz)
    v = (st["vx"] if axis == "x" else st["vz"]).reshape(nxx, nzz)
    w = st["vz"] if axis == "x" else st["vx"]
    line = (lambda a: a[:, 5]) if axis == "x" else (lambda a: a[4, :])
    assert rel_p(line(p)) < TOL
    assert rel_v(line(v)) < TOL
    # the field stays a function of one coordinate and the other velocity stays zero
    assert np.abs(w).max() < 1e-9 * vmax
    full = np.outer(p_cf, np.ones(nzz)) if axis == "x" else np.outer(np.ones(nxx), p_cf)
    assert np.linalg.norm(p - full) / (pmax * scale * np.sqrt(other)) < TOL


# ---------------------------------------------------------------------------
# every mode: homogeneous periodic medium vs the per-mode statement
# ---------------------------------------------------------------------------
GRIDS = [(24, 20), (25, 21), (27, 16), (18, 25)]


def ricker(nt, dt, fd=30.0):
    return host.ricker(nt, dt, fd, 2)


@pytest.mark.parametrize("nxx,nzz", GRIDS)
@pytest.mark.parametrize("absorb", [False, True])
def test_acoustic_all_modes(nxx, nzz, absorb):
    dt, nt, vp, rho = 1e-3, 40, 1800.0, 1024.0
    cp = dict(ia=-0.125, ib=0.75, ha=-0.25, hb=0.625) if absorb else None
    recs = ((3, 4), (nxx - 2, 1), (nxx // 2, nzz - 3))
    rik = ricker(nt, dt)
    job = homog_job(nxx, nzz, dt=dt, nt=nt, vp=vp, rho=rho, cpml=cp, src=(nxx // 3, nzz // 2 + 1),
                    rik=rik, recs=recs)
    rec, st = orc.forward(job, want_state=True)
    g, _, (p, vx, vz) = spectral.simulate(nxx, nzz, 10.0, 10.0, dt, nt, vp, rho, rik=rik.astype(float),
                                          src=(nxx // 3, nzz // 2 + 1), recs=recs, cpml=cp)
    assert rel(rec, g) < TOL_G
    for a, b in ((st["p"], p), (st["vx"], vx), (st["vz"], vz)):
        assert rel(a.reshape(nxx, nzz), b) < TOL


Q = dict(gam=0.0078125, lam=1.0, vwd=1.1875, tao=37.5)


@pytest.mark.parametrize("nxx,nzz", GRIDS[:3])
def test_visco_all_modes(nxx, nzz):
    dt, nt, vp, rho = 1e-3, 40, 1800.0, 1024.0
    cp = dict(ia=-0.125, ib=0.75, ha=-0.25, hb=0.625)
    recs = ((3, 4), (nxx - 2, 1))
    rik = ricker(nt, dt)
    q = dict(Q, lam=1.0625)
    job = homog_job(nxx, nzz, dt=dt, nt=nt, vp=vp, rho=rho, kind=1, cpml=cp, q=q, src=(4, 6), rik=rik,
                    recs=recs)
    rec, st = orc.forward(job, want_state=True)
    g, _, (p, vx, vz) = spectral.simulate(nxx, nzz, 10.0, 10.0, dt, nt, vp, rho, rik=rik.astype(float),
                                          src=(4, 6), recs=recs, cpml=cp, kind=1, q=q)
    assert rel(rec, g) < TOL_G
    assert rel(st["p"].reshape(nxx, nzz), p) < TOL
    # the loss terms matter at this size: the acoustic answer is far off
    ga, _, _ = spectral.simulate(nxx, nzz, 10.0, 10.0, dt, nt, vp, rho, rik=rik.astype(float), src=(4, 6),
                                 recs=recs, cpml=cp)
    assert rel(ga, g) > 1e-2


def test_q_to_infinity_equals_acoustic():
    """gamma = 0, lambda = 1, tau = 0 (Q -> infinity in transfermodel,
    func.c:912-929) turns sum_total into div - (dt^2 v^2 / 12) K and removes
    the loss terms: the viscoacoustic oracle equals the acoustic one, here on
    a layered model with CPML."""
    job = synth.small_job(nxt=40, nzt=24, pml=5, nt=120, kind=1, seed=11)
    n = job["nxx"] * job["nzz"]
    job["gam"] = np.zeros(n, np.float32)
    job["lam"] = np.ones(n, np.float32)
    job["tao"] = np.zeros(n, np.float32)
    a = orc.forward(dict(job, kind=0))
    v = orc.forward(job)
    assert rel(v, a) < 1e-12


def test_constant_q_plane_wave_decay():
    """A standing plane wave in a constant-Q medium (transfermodel of Q = 40,
    reference frequency at the wave's own frequency) loses amplitude as
    exp(-w t / (2 Q)); the oracle's loss terms give that rate within 15%
    (a dropped factor, a wrong sign or the wrong power of |k| does not)."""
    n, d, vp, rho, Qv = 32, 10.0, 2000.0, 1024.0, 40.0
    m = 4
    k = 2 * np.pi * m / (n * d)
    w = vp * k
    fd = w / (2 * np.pi)
    dt = 2e-4
    period = 2 * np.pi / w
    nt = int(round(6 * period / dt))
    qf = host.transfermodel(fd, fd, np.array([vp], np.float32), np.array([Qv], np.float32))
    q = {kk: float(qf[kk][0]) for kk in ("gam", "lam", "vwd", "tao")}
    job = homog_job(n, 8, dx=d, dz=d, dt=dt, nt=nt, vp=vp, rho=rho, kind=1, q=q, recs=((0, 0),))
    i = np.arange(n)
    p0 = np.outer(np.cos(2 * np.pi * m * i / n), np.ones(8))
    # energy-like amplitude: the p envelope over the run
    rec = orc.forward(job, init={"p": p0})[:, 0]
    t = np.arange(nt) * dt
    peaks = [j for j in range(1, nt - 1) if abs(rec[j]) >= abs(rec[j - 1]) and abs(rec[j]) >= abs(rec[j + 1])
             and abs(rec[j]) > 0.05]
    tp, ap = t[peaks], np.abs(rec[peaks])
    slope = np.polyfit(tp, np.log(ap), 1)[0]
    expect = -w / (2 * Qv)
    assert abs(slope / expect - 1) < 0.15, (slope, expect)


# ---------------------------------------------------------------------------
# RTM / Q-RTM
# ---------------------------------------------------------------------------
def rtm_spectral(nxx, nzz, pml, dt, nt, vp, rho, rik, src, cells, obs, cpml=None, q=None, etas=0.0,
                 etar=0.0):
    """Image and illumination of one shot by full storage: every source field
    of the forward sweep kept, the receiver wavefield stepped with data rows
    nt-1, nt-2, ..., 0 replacing p at the receiver cells, img = sum_t
    p_src(t) p_rcv(t) and ssg = sum_t p_src(t)^2 over the interior."""
    qs = -1.0 if q is not None else 1.0
    taoc = None if q is None else q["tao"]
    kw = dict(cpml=cpml, kind=1 if q is not None else 0, q=q, qsign=qs)
    _, fs, _ = spectral.simulate(nxx, nzz, 10.0, 10.0, dt, nt, vp, rho, rik=rik, src=src,
                                 stab=None if q is None else (etas, taoc), **kw)
    rows = [obs[nt - 1 - k] for k in range(nt)]
    _, fr, _ = spectral.simulate(nxx, nzz, 10.0, 10.0, dt, nt, vp, rho, replace=(cells, rows),
                                 stab=None if q is None else (etar, taoc), t_index=lambda k: nt - 1 - k,
                                 **kw)
    sl = (slice(pml, nxx - pml), slice(pml, nzz - pml))
    img = sum(fs[nt - 1 - k][sl] * fr[k][sl] for k in range(nt))
    ssg = sum(f[sl] ** 2 for f in fs)
    return img, ssg


@pytest.mark.parametrize("kind", [2, 3])
def test_rtm_all_modes_and_revolve(kind):
    nxx, nzz, pml, dt, nt, vp, rho = 24, 21, 3, 1e-3, 30, 1800.0, 1024.0
    cp = dict(ia=-0.125, ib=0.75, ha=-0.25, hb=0.625)
    rik = ricker(nt, dt).astype(np.float64)
    rik[0] = 0.0   # checkpoint 0 then holds the true state at time 0 (see below)
    cells = ((4, 5), (9, 5), (4, 5), (15, 6))   # a duplicate cell: the later receiver's data win
    rng = np.random.default_rng(5)
    obs = rng.standard_normal((nt, len(cells))).astype(np.float32)
    # a strong loss (tau) and stabilisers eta large enough that exp(-tau t k)
    # meets eta 1e-6 within the run, so that the stabiliser acts
    q = dict(Q, tao=500.0) if kind == 3 else None
    etas, etar = (2.0e4, 1.0e4) if kind == 3 else (0.0, 0.0)
    job = homog_job(nxx, nzz, dt=dt, nt=nt, vp=vp, rho=rho, kind=kind, cpml=cp, q=q, src=(11, 6), rik=rik,
                    recs=cells)
    job.update(pml=pml, obs=obs, etas=etas, etar=etar)
    img_s, ssg_s = rtm_spectral(nxx, nzz, pml, dt, nt, vp, rho, rik, (11, 6), cells, obs.astype(float),
                                cpml=cp, q=q, etas=etas, etar=etar)
    res = {}
    for snaps in (2, 5, nt):
        job["snaps"] = snaps
        res[snaps] = orc.rtm(job, host.revolve_trace(nt, snaps)[0])
        assert rel(res[snaps][0], img_s) < TOL, snaps
        assert rel(res[snaps][1], ssg_s) < TOL, snaps
    # revolve-reconstructed source fields equal stored ones exactly
    for snaps in (2, 5):
        assert np.array_equal(res[snaps][0], res[nt][0])
        assert np.array_equal(res[snaps][1], res[nt][1])
    if kind == 3:
        # the stabiliser acts at these settings (eta = 0 is the plain |k|^1.01)
        img0, _ = rtm_spectral(nxx, nzz, pml, dt, nt, vp, rho, rik, (11, 6), cells, obs.astype(float), cpml=cp,
                               q=q)
        assert rel(img0, img_s) > 1e-3
        job.update(etas=0.0, etar=0.0, snaps=5)
        assert rel(orc.rtm(job, host.revolve_trace(nt, 5)[0])[0], img0) < TOL


def test_rtm_checkpoint0_quirk_changes_reconstruction():
    """With rik[0] != 0 the reference's checkpoint 0 (zero state plus the first
    wavelet sample at the source, velocities not advanced,
    sgpstd-qrtm/acps4.cu:468-475, 1846-1863) differs from the true state after
    step 0, so an image whose source fields are rebuilt from it differs from
    full storage; with rik[0] = 0 it does not (previous test)."""
    nxx, nzz, pml, dt, nt = 24, 21, 3, 1e-3, 30
    rik = ricker(nt, dt).astype(np.float64)
    rik[0] = 0.5
    cells = ((4, 5), (15, 6))
    obs = np.random.default_rng(6).standard_normal((nt, 2)).astype(np.float32)
    job = homog_job(nxx, nzz, dt=dt, nt=nt, vp=1800.0, rho=1024.0, kind=2, src=(11, 6), rik=rik, recs=cells)
    job.update(pml=pml, obs=obs)
    job["snaps"] = nt
    full = orc.rtm(job, host.revolve_trace(nt, nt)[0])[0]
    job["snaps"] = 2
    rebuilt = orc.rtm(job, host.revolve_trace(nt, 2)[0])[0]
    assert rel(rebuilt, full) > 1e-6
