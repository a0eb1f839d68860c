"""ctypes front end of the CPU oracle (oracle/liboracle64.so, FP64) and of its
FP32 build (oracle/liboracle32.so).  Test infrastructure: used only as the
checker by tests/, smoke() and bench.py's CPU baseline."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_libs = {}


def lib(bits=64):
    if bits not in _libs:
        path = os.path.join(ORACLE_DIR, "liboracle%d.so" % bits)
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: %s (run make -C oracle)" % path)
        _libs[bits] = C.CDLL(path)
    return _libs[bits]


def _dt(bits):
    return np.float64 if bits == 64 else np.float32


def _ptr(a, bits):
    t = C.c_double if bits == 64 else C.c_float
    return a.ctypes.data_as(C.POINTER(t))


def fft1d(x, inverse=False):
    x = np.ascontiguousarray(x, np.complex128)
    out = np.zeros_like(x)
    lib(64).orc_fft1d(C.c_int(x.size), _ptr(x.view(np.float64), 64),
                      _ptr(out.view(np.float64), 64), C.c_int(1 if inverse else 0))
    return out


def fft2d(a, inverse=False):
    a = np.ascontiguousarray(a, np.complex128).copy()
    nxx, nzz = a.shape
    lib(64).orc_fft2d(C.c_int(nxx), C.c_int(nzz), _ptr(a.view(np.float64), 64),
                      C.c_int(1 if inverse else 0))
    return a


FIELD_KEYS = ("vp", "rho", "bbx", "bbz", "gam", "lam", "vwd", "tao")
C1D_KEYS = ("iax", "ibx", "iaz", "ibz", "hax", "hbx", "haz", "hbz",
            "kxsmr", "kxsmi", "kxscr", "kxsci", "kzsmr", "kzsmi", "kzscr", "kzsci")
K2_KEYS = ("k2d1", "k2d2", "k2d3", "k2da")


class _Args:
    """Converts a job dict (float32 arrays, the engine's inputs) to the
    oracle's precision and keeps the buffers alive for the call."""

    def __init__(self, job, bits):
        dt = _dt(bits)
        self.keep = []
        # the oracle dereferences these without checks: a missing one would
        # crash it, so refuse the job here
        q = int(job["kind"]) in (1, 3)
        need = FIELD_KEYS[:8 if q else 4] + C1D_KEYS + (K2_KEYS if q else K2_KEYS[:1])
        missing = [k for k in need if job.get(k) is None]
        if missing:
            raise ValueError("job lacks %s (build it with the k powers: want_k2=True)" % ", ".join(missing))
        PT = C.POINTER(C.c_double if bits == 64 else C.c_float)

        def arr(keys):
            ptrs = (PT * len(keys))()
            for i, k in enumerate(keys):
                if k in job and job[k] is not None:
                    a = np.ascontiguousarray(job[k], dt).ravel()
                    self.keep.append(a)
                    ptrs[i] = _ptr(a, bits)
            return ptrs

        self.fields = arr(FIELD_KEYS)
        self.c1d = arr(C1D_KEYS)
        self.k2 = arr(K2_KEYS)
        self.rik = np.ascontiguousarray(job["rik"], dt)
        self.gxg = np.ascontiguousarray(job["gxg"], np.int32)
        self.gzg = np.ascontiguousarray(job["gzg"], np.int32)


class lowrank_set:
    """Context: the oracle's viscoacoustic operator in rank-r low-rank form
    (U, A, V of cqh_lowrank) for the calls inside, as acps4.cu:1337-1459
    applies it; none (the Taylor form) otherwise."""

    def __init__(self, lr, bits=64):
        self.bits = bits
        dt = _dt(bits)
        self.arrs = None if lr is None else [np.ascontiguousarray(a, dt).ravel() for a in lr]
        self.rank = 0 if lr is None else int(np.asarray(lr[1]).shape[0])

    def __enter__(self):
        if self.arrs is not None:
            lib(self.bits).orc_set_lowrank(C.c_int(self.rank), *[_ptr(a, self.bits) for a in self.arrs])
        return self

    def __exit__(self, *a):
        lib(self.bits).orc_set_lowrank(C.c_int(0), None, None, None)


def forward(job, bits=64, want_state=False, init=None, lowrank=None):
    """One-shot forward modeling on the CPU.  job: dict with kind (0 acoustic,
    1 viscoacoustic), nxx, nzz, nt, dt and the float32 arrays of the job
    (window fields, 1-D coefficients, k powers, rik, gxg, gzg, xsg, zsg).
    init: optional {"p", "vx", "vz"} initial wavefield (nxx*nzz each)."""
    L = lib(bits)
    dt = _dt(bits)
    a = _Args(job, bits)
    nt, nr = int(job["nt"]), len(a.gxg)
    N = int(job["nxx"]) * int(job["nzz"])
    rec = np.zeros(nt * nr, dt)
    p = np.zeros(N, dt)
    vx = np.zeros(N, dt)
    vz = np.zeros(N, dt)
    real = C.c_double if bits == 64 else C.c_float
    ic = []
    for k in ("p", "vx", "vz"):
        if init is not None and init.get(k) is not None:
            v = np.ascontiguousarray(init[k], dt).ravel()
            assert v.size == N
            ic.append(v)
        else:
            ic.append(None)
    icp = [None if v is None else _ptr(v, bits) for v in ic]
    with lowrank_set(lowrank, bits):
        L.orc_forward_ic(C.c_int(int(job["kind"])), C.c_int(int(job["nxx"])), C.c_int(int(job["nzz"])),
                         C.c_int(nt), real(float(job["dt"])), a.fields, a.c1d, a.k2, _ptr(a.rik, bits),
                         C.c_int(int(job["xsg"])), C.c_int(int(job["zsg"])), C.c_int(nr),
                         a.gxg.ctypes.data_as(C.POINTER(C.c_int)),
                         a.gzg.ctypes.data_as(C.POINTER(C.c_int)), _ptr(rec, bits), _ptr(p, bits),
                         _ptr(vx, bits), _ptr(vz, bits), icp[0], icp[1], icp[2])
    rec = rec.reshape(nt, nr)
    if want_state:
        return rec, {"p": p, "vx": vx, "vz": vz}
    return rec


def rtm(job, trace, bits=64, lowrank=None):
    """RTM / Q-RTM of one shot (job kind 2 or 3, with obs, snaps, etas, etar)
    following the revolve action trace [(action, capo, check, fine), ...].
    Returns the raw interior (img, ssg), each (nx, nz)."""
    L = lib(bits)
    dt_ = _dt(bits)
    a = _Args(job, bits)
    nxx, nzz, pml = int(job["nxx"]), int(job["nzz"]), int(job["pml"])
    nx, nz = nxx - 2 * pml, nzz - 2 * pml
    obs = np.ascontiguousarray(job["obs"], dt_)
    act = np.array([x[0] for x in trace], np.int32)
    capo = np.array([x[1] for x in trace], np.int32)
    check = np.array([x[2] for x in trace], np.int32)
    img = np.zeros(nx * nz, dt_)
    ssg = np.zeros(nx * nz, dt_)
    real = C.c_double if bits == 64 else C.c_float
    ip = lambda v: v.ctypes.data_as(C.POINTER(C.c_int))
    with lowrank_set(lowrank, bits):
        _rtm_call(L, job, a, bits, nxx, nzz, pml, obs, act, capo, check, img, ssg, real, ip)
    return img.reshape(nx, nz), ssg.reshape(nx, nz)


def _rtm_call(L, job, a, bits, nxx, nzz, pml, obs, act, capo, check, img, ssg, real, ip):
    L.orc_rtm(C.c_int(int(job["kind"])), C.c_int(nxx), C.c_int(nzz), C.c_int(pml),
              C.c_int(int(job["nt"])), real(float(job["dt"])), a.fields, a.c1d, a.k2,
              _ptr(a.rik, bits), C.c_int(int(job["xsg"])), C.c_int(int(job["zsg"])),
              C.c_int(len(a.gxg)), ip(a.gxg), ip(a.gzg), _ptr(obs, bits), ip(act), ip(capo),
              ip(check), C.c_int(len(act)), C.c_int(int(job["snaps"])),
              real(float(job.get("etas", 0.0))), real(float(job.get("etar", 0.0))),
              _ptr(img, bits), _ptr(ssg, bits))


def set_threads(n, bits=64):
    lib(bits).orc_set_threads(C.c_int(n))


def num_threads(bits=64):
    return int(lib(bits).orc_num_threads())
