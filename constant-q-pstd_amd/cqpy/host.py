"""ctypes bindings for libcq_host.so (parameter files, coefficient builders,
shot windowing, revolve schedule).

Every builder returns float32 numpy arrays laid out like the reference
(x-major, z fastest: ``a[ix * nzz + iz]``).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)


def _fp(a):
    return a.ctypes.data_as(_F)


def _ip(a):
    return a.ctypes.data_as(_I)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(PKG_DIR, "libcq_host.so")
        if not os.path.exists(path):
            raise RuntimeError("libcq_host.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.cqh_check_stability.restype = C.c_float
        L.cqh_get_max_tao.restype = C.c_float
        L.cqh_window_sroffmx.restype = C.c_float
        L.cqh_lr_symbol.restype = C.c_double
        L.cqh_lr_symbol.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.cqh_lowrank_error.restype = C.c_double
        _lib = L
    return _lib


class ShotGeom(C.Structure):
    _fields_ = [("xsg", C.c_int), ("zsg", C.c_int), ("nr", C.c_int),
                ("lgest", C.c_int), ("rgest", C.c_int), ("nclamped", C.c_int),
                ("mlx", C.c_float)]


class Revolve(C.Structure):
    _fields_ = [("check", C.c_int), ("capo", C.c_int), ("fine", C.c_int),
                ("snaps", C.c_int), ("info", C.c_int), ("turn", C.c_int),
                ("oldsnaps", C.c_int), ("oldfine", C.c_int),
                ("ch", C.c_int * 64), ("advances", C.c_int),
                ("takeshots", C.c_int), ("commands", C.c_int)]


RV_NAMES = ["advance", "takeshot", "restore", "firsturn", "youturn", "terminate", "error"]


def pml_coeffs(pml, fm, nxx, nzz, dx, dz, dt):
    """get_pml_parameters (sgpstd-qmodeling/funca.c:223-340)."""
    out = {k: np.zeros(nxx, np.float32) for k in ("iax", "ibx", "hax", "hbx")}
    out.update({k: np.zeros(nzz, np.float32) for k in ("iaz", "ibz", "haz", "hbz")})
    lib().cqh_pml_coeffs(C.c_int(pml), C.c_float(fm), _fp(out["iax"]), _fp(out["ibx"]),
                         _fp(out["iaz"]), _fp(out["ibz"]), _fp(out["hax"]), _fp(out["hbx"]),
                         _fp(out["haz"]), _fp(out["hbz"]), C.c_int(nxx), C.c_int(nzz),
                         C.c_float(dx), C.c_float(dz), C.c_float(dt))
    return out


def shift_coeffs(dx, dz, nxx, nzz):
    """cal_shift_parameters (sgpstd-qmodeling/funca.c:342-427)."""
    o = {k: np.zeros(nxx, np.float32) for k in ("kxsmr", "kxscr", "kxsmi", "kxsci")}
    o.update({k: np.zeros(nzz, np.float32) for k in ("kzsmr", "kzscr", "kzsmi", "kzsci")})
    lib().cqh_shift_coeffs(_fp(o["kxsmr"]), _fp(o["kxscr"]), _fp(o["kzsmr"]), _fp(o["kzscr"]),
                           _fp(o["kxsmi"]), _fp(o["kxsci"]), _fp(o["kzsmi"]), _fp(o["kzsci"]),
                           C.c_float(dx), C.c_float(dz), C.c_int(nxx), C.c_int(nzz))
    return o


def k2d(dx, dz, nxx, nzz, q=True):
    """cal_k2d (funca.c:540-573 / func.c:550-583)."""
    n = nxx * nzz
    k1 = np.zeros(n, np.float32)
    if not q:
        lib().cqh_k2d(C.c_float(dx), C.c_float(dz), C.c_int(nxx), C.c_int(nzz), _fp(k1),
                      None, None, None)
        return {"k2d1": k1}
    k2, k3, ka = (np.zeros(n, np.float32) for _ in range(3))
    lib().cqh_k2d(C.c_float(dx), C.c_float(dz), C.c_int(nxx), C.c_int(nzz), _fp(k1), _fp(k2),
                  _fp(k3), _fp(ka))
    return {"k2d1": k1, "k2d2": k2, "k2d3": k3, "k2da": ka}


def bbx_bbz(nxx, nzz, rho):
    rho = np.ascontiguousarray(rho, np.float32).ravel()
    bx = np.zeros_like(rho)
    bz = np.zeros_like(rho)
    lib().cqh_bbx_bbz(C.c_int(nxx), C.c_int(nzz), _fp(rho), _fp(bx), _fp(bz))
    return bx, bz


def ricker(tmax, dt, fd, flag=2):
    r = np.zeros(tmax, np.float32)
    lib().cqh_ricker(C.c_int(tmax), C.c_float(dt), C.c_float(fd), _fp(r), C.c_int(flag))
    return r


def transfermodel(fd, fo, vp, qp):
    vp = np.ascontiguousarray(vp, np.float32).ravel()
    qp = np.ascontiguousarray(qp, np.float32).ravel()
    g, l, w, t = (np.zeros_like(vp) for _ in range(4))
    lib().cqh_transfermodel(C.c_int(vp.size), C.c_float(fd), C.c_float(fo), _fp(vp), _fp(qp),
                            _fp(g), _fp(l), _fp(w), _fp(t))
    return {"gam": g, "lam": l, "vwd": w, "tao": t}


def get_max_tao(tao):
    tao = np.ascontiguousarray(tao, np.float32).ravel()
    return float(lib().cqh_get_max_tao(C.c_int(tao.size), _fp(tao)))


def check_stability(vmax, dt, dx, dz):
    return float(lib().cqh_check_stability(C.c_float(vmax), C.c_float(dt), C.c_float(dx),
                                           C.c_float(dz)))


def window_field(full, nxt, nzt, nx, nz, pml, mlx, dx):
    """select_part_model + expand_pml_areas (funca.c:623-770)."""
    nxx, nzz = nx + 2 * pml, nz + 2 * pml
    full = np.ascontiguousarray(full, np.float32).ravel()
    win = np.zeros(nxx * nzz, np.float32)
    lib().cqh_select_part_model(C.c_int(pml), C.c_float(mlx), C.c_float(dx), C.c_int(nxt),
                                C.c_int(nzt), C.c_int(nx), C.c_int(nz), C.c_int(nxx),
                                C.c_int(nzz), _fp(full), _fp(win))
    lib().cqh_expand_pml(C.c_int(nx), C.c_int(nz), C.c_int(nxx), C.c_int(nzz), C.c_int(pml),
                         _fp(win))
    return win


def projectomodel(win_interior, nxt, nzt, nx, nz, pml, mlx, dx, full=None):
    if full is None:
        full = np.zeros(nxt * nzt, np.float32)
    win_interior = np.ascontiguousarray(win_interior, np.float32).ravel()
    lib().cqh_projectomodel(C.c_int(pml), C.c_float(mlx), C.c_float(dx), C.c_int(nxt),
                            C.c_int(nzt), C.c_int(nx), C.c_int(nz), C.c_int(nx + 2 * pml),
                            C.c_int(nz + 2 * pml), _fp(full), _fp(win_interior))
    return full


def filter_image(img, nx, nz, dx, dz):
    """5-point Laplacian of a raw interior image, zero border (sgpstd-qrtm/acps4.cu:1723-1745)."""
    img = np.ascontiguousarray(img, np.float32).ravel()
    out = np.zeros_like(img)
    lib().cqh_filter_image(C.c_int(nx), C.c_int(nz), C.c_float(dx), C.c_float(dz), _fp(img), _fp(out))
    return out


def mute_image(img, nx, nz, pml, lgest, rgest):
    """Zero interior columns outside the receiver span (sgpstd-qrtm/acps4.cu:926-935)."""
    img = np.ascontiguousarray(img, np.float32).ravel().copy()
    lib().cqh_mute_image(C.c_int(nx), C.c_int(nz), C.c_int(pml), C.c_int(lgest), C.c_int(rgest), _fp(img))
    return img


def removedirect(obs, tmax, nr, dt, fd, xs, zs, gxp, gzp, v0):
    obs = np.ascontiguousarray(obs, np.float32).ravel().copy()
    gxp = np.ascontiguousarray(gxp, np.float32)
    gzp = np.ascontiguousarray(gzp, np.float32)
    lib().cqh_removedirect(C.c_int(tmax), C.c_int(nr), C.c_float(dt), C.c_float(fd),
                           C.c_float(xs), C.c_float(zs), _fp(gxp), _fp(gzp), C.c_float(v0),
                           _fp(obs))
    return obs


def shot_indices(pml, nxx, nzz, dx, dz, sroffmx, disx, xs, zs, gxp, gzp):
    """Source / receiver grid indices (acps4.cu:391-426)."""
    gxp = np.ascontiguousarray(gxp, np.float32)
    gzp = np.ascontiguousarray(gzp, np.float32)
    nr = gxp.size
    gxg = np.zeros(nr, np.int32)
    gzg = np.zeros(nr, np.int32)
    g = ShotGeom()
    rc = lib().cqh_shot_indices(C.c_int(pml), C.c_int(nxx), C.c_int(nzz), C.c_float(dx),
                                C.c_float(dz), C.c_float(sroffmx), C.c_float(disx),
                                C.c_float(xs), C.c_float(zs), C.c_int(nr), _fp(gxp), _fp(gzp),
                                _ip(gxg), _ip(gzg), C.byref(g))
    if rc != 0:
        raise ValueError("shot grid index exceeds the window")
    return gxg, gzg, g


def numforw(steps, snaps):
    return int(lib().cqh_numforw(C.c_int(steps), C.c_int(snaps)))


def revolve_trace(steps, snaps, limit=10_000_000):
    """Full action trace [(action, capo, check, fine), ...] of the schedule."""
    r = Revolve()
    lib().cqh_revolve_init(C.byref(r), C.c_int(steps), C.c_int(snaps))
    out = []
    for _ in range(limit):
        a = lib().cqh_revolve_next(C.byref(r))
        out.append((a, r.capo, r.check, r.fine))
        if a in (5, 6):
            break
    return out, r.info


LR_TAYLOR4, LR_FRACLAP = 0, 1
_D = C.POINTER(C.c_double)


def lowrank_cc(vp, lam, dt):
    """cc = dt^2 v^2 lam / 12 per cell, in float32 as the engine forms it."""
    vp = np.asarray(vp, np.float32)
    lam = np.asarray(lam, np.float32)
    dtf = np.float32(dt) * np.float32(dt)
    return (np.float32(1.0 / 12.0) * dtf * (vp * vp) * lam).astype(np.float32)


def lowrank(job, rank, symbol=LR_FRACLAP, seed=1):
    """FP64 low-rank decomposition of the viscoacoustic operator of a job
    window (cqh_lowrank).  Returns (U [N1, r], A [r, r], V [r, Nq]) as
    float64 arrays."""
    nxx, nzz = int(job["nxx"]), int(job["nzz"])
    N1, Nq = nxx * nzz, (nxx // 2 + 1) * (nzz // 2 + 1)
    g = np.ascontiguousarray(job["gam"], np.float32).ravel()
    w = np.ascontiguousarray(job["vwd"], np.float32).ravel()
    cc = lowrank_cc(job["vp"], job["lam"], job["dt"]).ravel()
    U = np.zeros(N1 * rank)
    A = np.zeros(rank * rank)
    V = np.zeros(rank * Nq)
    rc = lib().cqh_lowrank(C.c_int(nxx), C.c_int(nzz), C.c_float(job["dx"]), C.c_float(job["dz"]), _fp(g), _fp(w),
                           _fp(cc), C.c_int(symbol), C.c_int(rank), C.c_ulonglong(seed),
                           U.ctypes.data_as(_D), A.ctypes.data_as(_D), V.ctypes.data_as(_D))
    if rc != 0:
        raise ValueError("low-rank decomposition failed")
    return U.reshape(N1, rank), A.reshape(rank, rank), V.reshape(rank, Nq)


def lowrank_pack(job, U, A, V):
    """Engine inputs (weights [r, N1], symbols / N [r, Nq]) as float32."""
    nxx, nzz = int(job["nxx"]), int(job["nzz"])
    r = A.shape[0]
    N1, Nq = nxx * nzz, (nxx // 2 + 1) * (nzz // 2 + 1)
    wts = np.zeros(r * N1, np.float32)
    sym = np.zeros(r * Nq, np.float32)
    lib().cqh_lowrank_pack(C.c_int(nxx), C.c_int(nzz), C.c_int(r), np.ascontiguousarray(U).ctypes.data_as(_D),
                           np.ascontiguousarray(A).ctypes.data_as(_D), np.ascontiguousarray(V).ctypes.data_as(_D),
                           _fp(wts), _fp(sym))
    return wts.reshape(r, N1), sym.reshape(r, Nq)


def lowrank_error(job, U, A, V, cells, symbol=LR_FRACLAP):
    """checka_error: max |W - U A V| / max |W| over the given cells."""
    nxx, nzz = int(job["nxx"]), int(job["nzz"])
    g = np.ascontiguousarray(job["gam"], np.float32).ravel()
    w = np.ascontiguousarray(job["vwd"], np.float32).ravel()
    cc = lowrank_cc(job["vp"], job["lam"], job["dt"]).ravel()
    cells = np.ascontiguousarray(cells, np.int32)
    return float(lib().cqh_lowrank_error(C.c_int(nxx), C.c_int(nzz), C.c_float(job["dx"]), C.c_float(job["dz"]),
                                         _fp(g), _fp(w), _fp(cc), C.c_int(symbol), C.c_int(A.shape[0]),
                                         np.ascontiguousarray(U).ctypes.data_as(_D),
                                         np.ascontiguousarray(A).ctypes.data_as(_D),
                                         np.ascontiguousarray(V).ctypes.data_as(_D), C.c_int(cells.size), _ip(cells)))


def lowrank_kgrid(nxx, nzz, dx, dz):
    k2 = np.zeros((nxx // 2 + 1) * (nzz // 2 + 1))
    lib().cqh_lr_kgrid(C.c_int(nxx), C.c_int(nzz), C.c_float(dx), C.c_float(dz), k2.ctypes.data_as(_D))
    return k2
