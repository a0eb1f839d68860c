"""Seeded synthetic models, geometry and shot jobs for the benchmark
workloads and the tests (there are no field data in this project).

Models are float32, x-major, z fastest (``m[ix * nzt + iz]``), the layout of
the reference's model files (README.md:55-57).  A job is a dict holding
exactly the float32 arrays the reference's host code would hand its engine for
one shot (window fields, CPML and shift coefficients, |k| powers, wavelet,
grid indices); it feeds both the HIP engine and the CPU oracle.
"""
import numpy as np

from . import host


def _smooth(a, sigma):
    """Separable Gaussian smoothing (reflect padding), numpy only."""
    if sigma <= 0:
        return a
    r = int(3 * sigma + 0.5)
    x = np.arange(-r, r + 1)
    g = np.exp(-0.5 * (x / sigma) ** 2)
    g /= g.sum()
    out = a.astype(np.float64)
    for ax in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[ax] = (r, r)
        b = np.pad(out, pad, mode="reflect")
        acc = np.zeros_like(out)
        for i, w in enumerate(g):
            sl = [slice(None), slice(None)]
            sl[ax] = slice(i, i + out.shape[ax])
            acc += w * b[tuple(sl)]
        out = acc
    return out


def layered_model(nxt, nzt, dx, seed=101, nlayers=6, vmin=1500.0, vmax=4000.0,
                  noise=0.03, noise_sigma=3.0, amp_m=40.0, qrange=None, anomaly=None,
                  salt=None, smooth_sigma=0.0):
    """Layered vp with sinusoidal interfaces and smoothed noise; Gardner
    density clamped to [1000, 2600]; optional per-layer Q with a low-Q
    Gaussian anomaly; optional high-velocity ellipse ("salt").  Returns
    (vp, rho, q) as float32 (nxt, nzt) arrays (q None if qrange is None)."""
    rng = np.random.default_rng(seed)
    depths = np.sort(rng.uniform(0.08, 0.92, nlayers - 1)) * nzt
    x = np.arange(nxt) * dx
    vel = np.linspace(vmin, vmax, nlayers)
    qv = None
    if qrange is not None:
        qv = rng.uniform(qrange[0], qrange[1], nlayers)
    iz = np.arange(nzt)[None, :].repeat(nxt, 0).astype(np.float64)
    layer = np.zeros((nxt, nzt), np.int32)
    for k, d in enumerate(depths):
        amp = rng.uniform(0.3, 1.0) * amp_m / dx
        lam = rng.uniform(0.3, 1.0) * nxt * dx
        ph = rng.uniform(0, 2 * np.pi)
        surf = d + amp * np.sin(2 * np.pi * x / lam + ph)
        layer += (iz >= surf[:, None]).astype(np.int32)
    vp = vel[layer]
    nz = _smooth(rng.standard_normal((nxt, nzt)), noise_sigma)
    nz = nz / (np.abs(nz).max() + 1e-30) * noise
    vp = vp * (1.0 + nz)
    if salt is not None:
        cx, cz, ax_, az_, vs = salt
        xx = np.arange(nxt)[:, None]
        zz = np.arange(nzt)[None, :]
        vp = np.where(((xx - cx) / ax_) ** 2 + ((zz - cz) / az_) ** 2 <= 1.0, vs, vp)
    if smooth_sigma > 0:
        vp = _smooth(vp, smooth_sigma)
    rho = np.clip(310.0 * vp ** 0.25, 1000.0, 2600.0)
    q = None
    if qv is not None:
        q = qv[layer].astype(np.float64)
        if anomaly is not None:
            cx, cz, rad, qa = anomaly
            xx = np.arange(nxt)[:, None]
            zz = np.arange(nzt)[None, :]
            w = np.exp(-(((xx - cx) ** 2 + (zz - cz) ** 2) / (rad ** 2)))
            q = q * (1 - w) + qa * w
        q = q.astype(np.float32)
    return vp.astype(np.float32), rho.astype(np.float32), q


def make_job(kind, vpt, rhot, nxt, nzt, dx, dz, dt, nt, fd, xs, zs, gxp, gzp, pml,
             qt=None, fo=None, disx=0.0, sroffmx=None, flagsource=2, want_k2=True):
    """Builds one shot the way the reference's driver + acforward do
    (acousticsgps4.c:239-414, acps4.cu:330-426; Q: visacousticsgps4.c:242,
    acsglw.cu:330-360).  kind 0 acoustic, 1 viscoacoustic."""
    vpt = np.ascontiguousarray(vpt, np.float32).ravel()
    rhot = np.ascontiguousarray(rhot, np.float32).ravel()
    gxp = np.ascontiguousarray(gxp, np.float32)
    gzp = np.ascontiguousarray(gzp, np.float32)
    if sroffmx is None:
        xr, xl = float(gxp.max()), float(gxp.min())
        sroffmx = np.float32(max(abs(np.float32(xr) - np.float32(xs)),
                                 abs(np.float32(xs) - np.float32(xl))))
    simxlen = np.float32(2.0 * (np.float32(sroffmx) + np.float32(disx)))
    nx = int(np.float64(np.float32(simxlen / np.float32(dx))) + 1.5)
    nz = nzt
    nxx, nzz = nx + 2 * pml, nz + 2 * pml
    bbxt, bbzt = host.bbx_bbz(nxt, nzt, rhot)
    mlx = np.float32(np.float32(xs) - np.float32(sroffmx) - np.float32(disx))
    win = lambda f: host.window_field(f, nxt, nzt, nx, nz, pml, float(mlx), dx)
    job = dict(kind=kind, nxx=nxx, nzz=nzz, nx=nx, nz=nz, pml=pml, nt=nt, dt=dt, dx=dx, dz=dz,
               mlx=float(mlx), sroffmx=float(sroffmx), disx=disx, fd=fd)
    job["vp"], job["rho"], job["bbx"], job["bbz"] = win(vpt), win(rhot), win(bbxt), win(bbzt)
    if kind in (1, 3):
        tq = host.transfermodel(fd, fo, vpt, qt)
        for k in ("gam", "lam", "vwd", "tao"):
            job[k] = win(tq[k])
    job.update(host.pml_coeffs(pml, fd, nxx, nzz, dx, dz, dt))
    job.update(host.shift_coeffs(dx, dz, nxx, nzz))
    if want_k2:
        job.update(host.k2d(dx, dz, nxx, nzz, q=(kind in (1, 3))))
    job["rik"] = host.ricker(nt, dt, fd, flagsource)
    gxg, gzg, geo = host.shot_indices(pml, nxx, nzz, dx, dz, float(sroffmx), disx, xs, zs,
                                      gxp, gzp)
    job.update(gxg=gxg, gzg=gzg, xsg=geo.xsg, zsg=geo.zsg, lgest=geo.lgest, rgest=geo.rgest)
    return job


def config1_model(seed=101):
    """160 x 398 layered model of the plain acoustic workload."""
    return layered_model(398, 160, 10.0, seed=seed)


def config1_job(nt=3000, kind=0, want_k2=True):
    """Plain acoustic one-shot workload: 160x398 interior, 30-point CPML,
    dx=dz=10 m, dt=0.5 ms, 15 Hz Ricker at x=199 dx, z=2 dz, 398 receivers
    at z=2 dz (padded window 459 x 220 by the reference's window rule).
    kind=1 gives the viscoacoustic variant (Q per layer 20-200 plus a Q=15
    Gaussian anomaly of radius 30 cells, fo = fd = 15 Hz)."""
    vp, rho, q = layered_model(398, 160, 10.0, seed=101, qrange=(20.0, 200.0),
                               anomaly=(199, 100, 30, 15.0))
    dx = 10.0
    gxp = (np.arange(398) * dx).astype(np.float32)
    gzp = np.full(398, 2 * dx, np.float32)
    return make_job(kind, vp, rho, 398, 160, dx, dx, 0.0005, nt, 15.0, 199 * dx, 2 * dx,
                    gxp, gzp, 30, qt=q, fo=15.0, want_k2=want_k2)


def small_job(nxt=40, nzt=24, pml=5, nt=60, kind=0, seed=7, dx=10.0, dt=0.001, fd=25.0,
              fo=None, src_col=None, nrec=None):
    """Small seeded shot for parity tests."""
    vp, rho, q = layered_model(nxt, nzt, dx, seed=seed, nlayers=3, qrange=(20.0, 200.0))
    src_col = nxt // 2 if src_col is None else src_col
    nrec = nxt if nrec is None else nrec
    gxp = (np.linspace(0, nxt - 1, nrec) * dx).astype(np.float32)
    gzp = np.full(nrec, 2 * dx, np.float32)
    return make_job(kind, vp, rho, nxt, nzt, dx, dx, dt, nt, fd, src_col * dx, 2 * dx, gxp,
                    gzp, pml, qt=q, fo=(fo if fo is not None else 2 * fd))


SUPPORTED_PRIMES = (2, 3, 5, 7, 11, 13, 17)


def fft_friendly(n, primes=(2, 3, 5, 7)):
    """Smallest m >= n whose prime factors are all in `primes`."""
    m = n
    while True:
        k = m
        for p in primes:
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def natural_size(job, nxx, nzz):
    """Record the padded window a request asked for (nxx_natural x
    nzz_natural cells, CPML included) beside the FFT-friendly window it is
    computed on: throughput counts the requested cells, not the widening."""
    job["nxx_natural"], job["nzz_natural"] = int(nxx), int(nzz)
    return job


def useful_cells(job):
    """Cells of the window a request asked for (see natural_size)."""
    return job.get("nxx_natural", job["nxx"]) * job.get("nzz_natural", job["nzz"])


def make_window_job(kind, vpt, rhot, nxt, nzt, dx, dz, dt, nt, fd, left_col, nx, src_col,
                    src_z, rec_cols, rec_z, pml, qt=None, fo=None, flagsource=2, want_k2=False):
    """A shot on an explicit window of model columns [left_col, left_col+nx)
    (columns outside the model replicate its edge, as select_part_model does,
    funca.c:623-733).  Used by the batched workload, whose windows are
    cropped around each source and therefore uneven."""
    vpt = np.ascontiguousarray(vpt, np.float32).ravel()
    rhot = np.ascontiguousarray(rhot, np.float32).ravel()
    nz = nzt
    nxx, nzz = nx + 2 * pml, nz + 2 * pml
    mlx = float(np.float32(left_col * dx))
    bbxt, bbzt = host.bbx_bbz(nxt, nzt, rhot)
    win = lambda f: host.window_field(f, nxt, nzt, nx, nz, pml, mlx, dx)
    job = dict(kind=kind, nxx=nxx, nzz=nzz, nx=nx, nz=nz, pml=pml, nt=nt, dt=dt, dx=dx, dz=dz,
               mlx=mlx, fd=fd)
    job["vp"], job["rho"], job["bbx"], job["bbz"] = win(vpt), win(rhot), win(bbxt), win(bbzt)
    if kind in (1, 3):
        tq = host.transfermodel(fd, fo, vpt, qt)
        for k in ("gam", "lam", "vwd", "tao"):
            job[k] = win(tq[k])
    job.update(host.pml_coeffs(pml, fd, nxx, nzz, dx, dz, dt))
    job.update(host.shift_coeffs(dx, dz, nxx, nzz))
    if want_k2:
        job.update(host.k2d(dx, dz, nxx, nzz, q=(kind in (1, 3))))
    job["rik"] = host.ricker(nt, dt, fd, flagsource)
    rec_cols = np.asarray(rec_cols)
    job["gxg"] = np.clip(rec_cols - left_col + pml, pml, nxx - pml - 1).astype(np.int32)
    job["gzg"] = np.full(rec_cols.size, rec_z + pml, np.int32)
    job["xsg"], job["zsg"] = int(src_col - left_col + pml), int(src_z + pml)
    return job


def batch128_jobs(nt=5000, nshots=128, kinds=None, want_k2=False, shots=None, widen=True, depth_pad=0):
    """Batched multi-shot workload: 128 shots on a 600 x 2400 model, 40-point
    CPML, dx = dz = 5 m, dt = 0.4 ms, layered vp 1500-4500 m/s with a salt-like
    ellipse at 4500 m/s, Q 30-300, 20 Hz Ricker, shots every 18 cells at the
    surface, receivers on every surface column inside each shot's window.
    Windows are cropped to +-600 cells around the source (uneven near the
    model edges; interior shots 1281 x 680 padded, 1281 = 3 7 61) and, with
    widen, widened to the right to the next 7-smooth padded width (1296).
    depth_pad extends the model below its last row (replicated, as the
    window rule replicates the model edge) so that the padded depth
    600 + depth_pad + 80 has no large prime factor.
    kinds[i] = 0 (acoustic) or 1 (viscoacoustic)."""
    nxt, nzt, dx, pml = 2400, 600, 5.0, 40
    vp, rho, q = layered_model(nxt, nzt, dx, seed=301, nlayers=6, vmin=1500.0, vmax=4500.0,
                               qrange=(30.0, 300.0), salt=(1200, 330, 260, 90, 4500.0))
    if depth_pad:
        ext = lambda a: np.concatenate([a, np.repeat(a[:, -1:], depth_pad, axis=1)], axis=1)
        vp, rho, q = ext(vp), ext(rho), ext(q)
        nzt += depth_pad
    first = (nxt - 1 - 18 * (nshots - 1)) // 2
    jobs = []
    for i in range(nshots) if shots is None else shots:
        c = first + 18 * i
        left = max(0, c - 600)
        right = min(nxt - 1, c + 600)
        nx = nx0 = right - left + 1
        if widen:
            nx = fft_friendly(nx + 2 * pml) - 2 * pml
        kind = 0 if kinds is None else kinds[i]
        jobs.append(make_window_job(kind, vp, rho, nxt, nzt, dx, dx, 0.0004, nt, 20.0, left, nx, c,
                                    2, np.arange(left, right + 1), 2, pml, qt=q, fo=20.0,
                                    want_k2=want_k2))
        natural_size(jobs[-1], nx0 + 2 * pml, 600 + 2 * pml)
    return jobs


SERVE_CLASSES = ((0, 0.5), (1, 0.3), (3, 0.2))   # acoustic, viscoacoustic, Q-RTM


def serve_jobs(rate=40.0, window_s=1.0, seed=501, nx_range=(512, 4096), nz_range=(256, 1024),
               nt_range=(2000, 8000), dx_range=(5.0, 12.5), fd_range=(10.0, 30.0), pml=40, snaps=40):
    """Serving workload: Poisson arrivals (mean `rate` jobs/s) over `window_s`
    seconds of mixed requests, 50% acoustic forward, 30% viscoacoustic forward,
    20% Q-RTM single shot.  Each request draws its window (nx columns, nz
    rows, widened to FFT-friendly padded sizes), nt, dx = dz, Ricker peak
    frequency and a layered vp/rho/Q model (vp 1500-4500 m/s, Q 30-300) from
    its own seed; dt = 0.4 dx / vmax.  The source sits at the window centre,
    receivers on every surface column.  Q-RTM requests carry seeded random
    observed data (the timing does not depend on its content) and `snaps`
    checkpoints.  Returns [(arrival seconds, job dict, class)]."""
    rng = np.random.default_rng(seed)
    out = []
    t = rng.exponential(1.0 / rate)
    k = 0
    while t < window_s:
        r = np.random.default_rng(seed * 1000 + k)
        u = r.uniform()
        kind = SERVE_CLASSES[0][0] if u < 0.5 else (SERVE_CLASSES[1][0] if u < 0.8 else SERVE_CLASSES[2][0])
        nx0 = int(r.integers(nx_range[0], nx_range[1] + 1))
        nz0 = int(r.integers(nz_range[0], nz_range[1] + 1))
        nx = fft_friendly(nx0 + 2 * pml) - 2 * pml
        nz = fft_friendly(nz0 + 2 * pml) - 2 * pml
        nt = int(r.integers(nt_range[0], nt_range[1] + 1))
        dx = float(np.round(r.uniform(*dx_range) * 4) / 4)
        fd = float(np.round(r.uniform(*fd_range)))
        vmax = 4500.0
        dt = float(np.float32(0.4 * dx / vmax))
        vp, rho, q = layered_model(nx, nz, dx, seed=seed * 1000 + k, vmin=1500.0, vmax=vmax,
                                   qrange=(30.0, 300.0))
        job = make_window_job(kind, vp, rho, nx, nz, dx, dx, dt, nt, fd, 0, nx, nx // 2, 2,
                              np.arange(nx), 2, pml, qt=q, fo=fd)
        natural_size(job, nx0 + 2 * pml, nz0 + 2 * pml)
        if kind == 3:
            job["obs"] = r.standard_normal((nt, len(job["gxg"]))).astype(np.float32)
            job["snaps"] = snaps
            job["etas"] = job["etar"] = 1e-20
        out.append((float(t), job, kind))
        t += rng.exponential(1.0 / rate)
        k += 1
    return out


def qrtm_jobs(nshots=32, nt=6000, snaps=40, kind=3, seed=401, obs_seed=402, widen=True):
    """Q-RTM workload: shots over an 800 x 3000 model with a 40-point CPML,
    dx = dz = 6.25 m, dt = 0.5 ms, 20 Hz Ricker, receivers +-1000 cells
    around each source at the surface (2001 columns, padded 2081, a prime;
    with widen the window is widened to 2020 columns so that the padded width
    2100 = 2^2 3 5^2 7 is 7-smooth), `snaps` checkpoints,
    stabiliser 1e-20 on both wavefields.  The migration model is a layered
    vp 1500-4500 m/s / Q 25-250 model smoothed with a Gaussian of 12 cells;
    observed data are seeded random traces (the run time does not depend on
    their content).  kind 3 = Q-RTM, 2 = acoustic RTM."""
    nxt, nzt, dx, pml = 3000, 800, 6.25, 40
    vp, rho, q = layered_model(nxt, nzt, dx, seed=seed, nlayers=8, vmin=1500.0, vmax=4500.0,
                               qrange=(25.0, 250.0), smooth_sigma=12.0)
    nx = fft_friendly(2001 + 2 * pml) - 2 * pml if widen else 2001
    first, last = 1000, nxt - 1001
    rng = np.random.default_rng(obs_seed)
    jobs = []
    for i in range(nshots):
        c = first + (last - first) * i // max(1, nshots - 1)
        left = c - nx // 2
        recs = np.arange(max(0, c - 1000), min(nxt, c + 1001))
        job = make_window_job(kind, vp, rho, nxt, nzt, dx, dx, 0.0005, nt, 20.0, left, nx, c, 2,
                              recs, 2, pml, qt=q, fo=20.0)
        job["obs"] = rng.standard_normal((nt, len(job["gxg"]))).astype(np.float32)
        job["snaps"] = snaps
        job["etas"] = job["etar"] = 1e-20
        jobs.append(natural_size(job, 2001 + 2 * pml, nzt + 2 * pml))
    return jobs


def with_lowrank(job, rank=4, symbol=None, seed=1):
    """The job with its viscoacoustic operator in rank-r low-rank form (FP64
    host decomposition, cqh_lowrank; default symbol: the fractional-Laplacian
    operator).  Returns (job, (U, A, V)); the decomposition feeds the
    oracle's low-rank apply."""
    sym = host.LR_FRACLAP if symbol is None else symbol
    U, A, V = host.lowrank(job, rank, symbol=sym, seed=seed)
    w, s = host.lowrank_pack(job, U, A, V)
    out = dict(job, lr_rank=rank, lr_w=w, lr_sym=s)
    return out, (U, A, V)
