"""An independent statement of the reference's time step for a HOMOGENEOUS,
PERIODIC medium, used only to check the oracle (test infrastructure).

The oracle (oracle/cq_oracle.c) follows the reference kernel by kernel: full
2-D complex FFTs of the padded grid, complex multipliers selected by index,
real parts taken after every inverse transform.  Here the same step is
written per Fourier mode instead.  With uniform coefficients every operator
of the step is diagonal in the wavenumber, so the state is advanced as the
2-D spectra (numpy's FFT) of the real fields, each mode on its own:

* A staggered derivative whose reference multiplier is m(k) acts on a real
  field as the Hermitian part h(k) = (m(k) + conj m(-k)) / 2 (taking the real
  part of the inverse transform of m F f).  m(k) is derived from the
  reference's table rule (sgpstd-qmodeling/funca.c:342-427 with the index
  selection of acps4.cu:1047-1108, 1210-1271), in closed form:
      v -> p : m(k) =  i k exp(-i k d / 2)   for index i <= n/2
               m(k) = -i |k| exp(+i theta')   for i > n/2, theta' = pi (n/2*2 - i) / n
      p -> v : the same with the opposite phase signs
  theta' is the phase of the mirror index the reference uses; for odd n it
  is one index short of |k| d / 2 (the odd-n quirk), so h is not i k e^{..}.
* Constant CPML coefficients a, b make the memory variables
  psi <- b psi + a d, d <- d + psi per-mode recurrences too.
* Sources (p[src] += rik[t] dt 1000, acps4.cu:1491) are the spectrum of a
  point; receivers read the inverse transform; RTM data replace p at the
  receiver cells (addsa_b_kernel, sgpstd-qrtm/acps4.cu:1662-1692).

The wavenumber grid follows the reference's float32 tables: dk = 2 pi/(n d)
with pi = 3.1415926 (funca.c:6), |k| mirrored above n/2.
"""
import numpy as np

PI_REF = 3.1415926


def kgrid(n, d):
    """|k| of each FFT index as the reference's float32 tables hold it."""
    dk = np.float32(2 * PI_REF / (n * d))
    i = np.arange(n)
    mir = np.where(i <= n // 2, i, n - i)
    return (mir.astype(np.float32) * dk).astype(np.float64)


def half_shift(n, d, sign):
    """Hermitian part of the staggered derivative multiplier along one axis:
    sign -1 for the v -> p derivative (phase e^{-i k d/2}), +1 for p -> v."""
    k = kgrid(n, d)
    i = np.arange(n)
    j = np.where(i <= n // 2, i, (n // 2) * 2 - i)          # index whose phase is used
    theta = (j * PI_REF / n).astype(np.float32).astype(np.float64)
    s = np.where(i <= n // 2, 1.0, -1.0)                     # sign of the true wavenumber
    m = s * 1j * k * np.exp(1j * sign * s * theta)
    return 0.5 * (m + np.conj(m[(-i) % n]))


def point_spectrum(nxx, nzz, ix, iz, amp=1.0):
    a = np.zeros((nxx, nzz))
    a[ix, iz] = amp
    return np.fft.fft2(a)


def simulate(nxx, nzz, dx, dz, dt, nt, vp, rho, rik=None, src=None, recs=(), kind=0,
             cpml=None, init=None, q=None, qsign=1.0, stab=None, replace=None, t_index=None, W=None):
    """Advances the homogeneous periodic medium nt steps.

    kind 0 acoustic (acps4.cu:508-607), 1 viscoacoustic (acsglw.cu:511-645).
    cpml: dict of scalar ia, ib (v -> p) and ha, hb (p -> v), the same on both
    axes; None for no absorption.  init: dict of initial real fields p, vx, vz.
    q: dict of scalar gam, lam, vwd, tao.  qsign -1 flips the loss terms
    (Q-compensated back-propagation, sgpstd-qrtm/acsglw.cu:1769-1806).
    W: optional function of k2 replacing the Taylor operator applied to div
    (the low-rank path's symbol).
    stab: (eta, taoc) for the stabilised aux multiplier of Q-RTM, with the
    step's time index from t_index(t) (default t).  replace: (cells, values)
    with values[t][j] written into p at cells[j] after the p update (RTM
    receiver wavefield; later cells win).
    Returns (gather nt x len(recs), list of p fields after each step, final
    (p, vx, vz))."""
    kx, kz = kgrid(nxx, dx), kgrid(nzz, dz)
    hxm, hxp = half_shift(nxx, dx, -1)[:, None], half_shift(nxx, dx, +1)[:, None]
    hzm, hzp = half_shift(nzz, dz, -1)[None, :], half_shift(nzz, dz, +1)[None, :]
    k2 = kx[:, None] ** 2 + kz[None, :] ** 2
    z = lambda: np.zeros((nxx, nzz), complex)
    P, X, Z = z(), z(), z()
    if init is not None:
        P = np.fft.fft2(init.get("p", np.zeros((nxx, nzz))))
        X = np.fft.fft2(init.get("vx", np.zeros((nxx, nzz))))
        Z = np.fft.fft2(init.get("vz", np.zeros((nxx, nzz))))
    psx, psz, phx, phz, aux = z(), z(), z(), z(), z()
    srcspec = None if src is None else point_spectrum(nxx, nzz, src[0], src[1])
    bb = 1.0 / rho                                  # cal_bbx_bbz of a uniform density
    c = rho * vp * vp * dt
    gather = np.zeros((nt, len(recs)))
    fields = []
    if q is not None:
        g, lam, vwd, tao = q["gam"], q["lam"], q["vwd"], q["tao"]
        s2 = k2 ** (1.0 / 32.0)                     # |k|^(1/16)
        ka = k2 ** 0.505                            # |k|^1.01
        # the mixed-domain operator applied to div (sum_total, acsglw.cu:1213-1245)
        if W is not None:
            W = W(k2)
        else:
            W = (1 - 32 * g) + 32 * g * vwd * s2 - dt * dt * vp * vp * lam / 12.0 * (
                (1 - 64 * g) * k2 + 64 * g * vwd * k2 * s2)
    for t in range(nt):
        dxv, dzv = hxm * X, hzm * Z
        if cpml is not None:
            psx = cpml["ib"] * psx + cpml["ia"] * dxv
            psz = cpml["ib"] * psz + cpml["ia"] * dzv
            dxv, dzv = dxv + psx, dzv + psz
        div = dxv + dzv
        if q is None:
            P = P + c * (div - (vp * dt) ** 2 / 12.0 * k2 * div)
        else:
            tot = W * div
            P = P + c * lam * (tot - qsign * 0.5 * dt * tao * ka * tot) - qsign * dt * tao * aux
        if srcspec is not None:
            P = P + rik[t] * dt * 1000.0 * srcspec
        if replace is not None or recs:
            p = np.fft.ifft2(P).real
            if replace is not None:
                cells, vals = replace
                for j, (ix, iz) in enumerate(cells):
                    p[ix, iz] = vals[t][j]
                P = np.fft.fft2(p)
            for j, (ix, iz) in enumerate(recs):
                gather[t, j] = p[ix, iz]
        fields.append(np.fft.ifft2(P).real)
        dxp, dzp = hxp * P, hzp * P
        if cpml is not None:
            phx = cpml["hb"] * phx + cpml["ha"] * dxp
            phz = cpml["hb"] * phz + cpml["ha"] * dzp
            dxp, dzp = dxp + phx, dzp + phz
        X = X + dt * bb * dxp
        Z = Z + dt * bb * dzp
        if q is not None:
            m = ka
            if stab is not None:
                eta, taoc = stab
                e = eta * 1e-6
                tt = (t if t_index is None else t_index(t)) * dt
                # multt_kernel, sgpstd-qrtm/acsglw.cu:1895-1924
                m = ka - 2 * e * ka * (1 + taoc * dt * ka) / (np.exp(-taoc * tt * ka) + e * np.exp(taoc * dt * ka))
            aux = m * P
    final = tuple(np.fft.ifft2(A).real for A in (P, X, Z))
    return gather, fields, final
