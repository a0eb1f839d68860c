"""ctypes bindings of libcq_hip.so (include/cq.h).

The product path: every call goes to the HIP engine; if the library or a GPU
is missing the calls raise instead of falling back to anything else.
"""
import ctypes as C
import os
import weakref

import numpy as np

from .host import PKG_DIR

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)


class JobDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("nxx", C.c_int), ("nzz", C.c_int), ("pml", C.c_int),
        ("nt", C.c_int), ("dt", C.c_float),
        ("vp", _F), ("rho", _F), ("bbx", _F), ("bbz", _F),
        ("gam", _F), ("lam", _F), ("vwd", _F), ("tao", _F),
        ("iax", _F), ("ibx", _F), ("hax", _F), ("hbx", _F),
        ("iaz", _F), ("ibz", _F), ("haz", _F), ("hbz", _F),
        ("kxsmr", _F), ("kxsmi", _F), ("kxscr", _F), ("kxsci", _F),
        ("kzsmr", _F), ("kzsmi", _F), ("kzscr", _F), ("kzsci", _F),
        ("dx", C.c_float), ("dz", C.c_float),
        ("rik", _F), ("xsg", C.c_int), ("zsg", C.c_int),
        ("nr", C.c_int), ("gxg", _I), ("gzg", _I),
        ("obs", _F), ("snaps", C.c_int), ("etas", C.c_float), ("etar", C.c_float),
        ("lr_rank", C.c_int), ("lr_w", _F), ("lr_sym", _F),
    ]


class ServerDesc(C.Structure):
    _fields_ = [("max_nxx", C.c_int), ("max_nzz", C.c_int), ("hbm_budget", C.c_longlong),
                ("max_jobs", C.c_int), ("run_cells", C.c_longlong)]


class Tiles(C.Structure):
    """cq_tiles: workgroup tile shapes of the passes and the FFT plan variant."""
    _fields_ = [("tz", C.c_int), ("cw", C.c_int), ("cwq", C.c_int), ("tk", C.c_int), ("tk2", C.c_int),
                ("plan", C.c_int)]

    @classmethod
    def of(cls, d):
        """From a dict (missing fields: the engine's choice)."""
        return cls(*(int(d.get(k, 0)) for k in ("tz", "cw", "cwq", "tk", "tk2")), int(d.get("plan", -1)))

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in ("tz", "cw", "cwq", "tk", "tk2", "plan")}


QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3
NPASS = 10   # CQ_NPASS

SYMBOLS = ["cq_open", "cq_close", "cq_last_error", "cq_version", "cq_batch_create",
           "cq_batch_step", "cq_batch_sync", "cq_batch_run", "cq_batch_read_rec",
           "cq_batch_read_p", "cq_batch_time", "cq_batch_bytes", "cq_batch_destroy",
           "cq_plain_forward", "cq_batch_enable_timing", "cq_batch_pass_times",
           "cq_copy_calibrate", "cq_batch_kernel_times", "cq_pass_name", "cq_batch_read_image",
           "cq_server_open", "cq_server_submit", "cq_server_advance", "cq_server_poll",
           "cq_server_job_steps", "cq_server_result", "cq_server_release", "cq_server_close",
           "cq_forward_static", "cq_batch_create_tiled", "cq_batch_get_tiles", "cq_autotune",
           "cq_batch_use_graphs"]

_lib = None


def lib_path():
    return os.path.join(PKG_DIR, "libcq_hip.so")


def lib():
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("libcq_hip.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.cq_last_error.restype = C.c_char_p
        L.cq_version.restype = C.c_char_p
        L.cq_batch_bytes.restype = C.c_longlong
        L.cq_pass_name.restype = C.c_char_p
        _lib = L
    return _lib


class CqError(RuntimeError):
    pass


def _chk(rc):
    if rc != 0:
        raise CqError("cq error %d: %s" % (rc, lib().cq_last_error().decode()))


_F32_KEYS = ("vp", "rho", "bbx", "bbz", "gam", "lam", "vwd", "tao", "iax", "ibx", "hax", "hbx",
             "iaz", "ibz", "haz", "hbz", "kxsmr", "kxsmi", "kxscr", "kxsci", "kzsmr", "kzsmi",
             "kzscr", "kzsci", "rik")


def make_desc(job, keep):
    """cq_job_desc from a job dict (see cqpy.synth.make_job); `keep` collects
    the arrays that must stay alive until the descriptor is consumed."""
    d = JobDesc()
    d.kind, d.nxx, d.nzz, d.pml = int(job["kind"]), int(job["nxx"]), int(job["nzz"]), int(job["pml"])
    d.nt, d.dt = int(job["nt"]), float(job["dt"])
    d.dx, d.dz = float(job["dx"]), float(job["dz"])
    for k in _F32_KEYS:
        if k in job and job[k] is not None:
            a = np.ascontiguousarray(job[k], np.float32)
            keep.append(a)
            setattr(d, k, a.ctypes.data_as(_F))
    gx = np.ascontiguousarray(job["gxg"], np.int32)
    gz = np.ascontiguousarray(job["gzg"], np.int32)
    keep += [gx, gz]
    d.xsg, d.zsg, d.nr = int(job["xsg"]), int(job["zsg"]), gx.size
    d.gxg, d.gzg = gx.ctypes.data_as(_I), gz.ctypes.data_as(_I)
    if job.get("obs") is not None:
        ob = np.ascontiguousarray(job["obs"], np.float32)
        keep.append(ob)
        d.obs = ob.ctypes.data_as(_F)
    if int(job.get("lr_rank", 0)) > 0:
        w = np.ascontiguousarray(job["lr_w"], np.float32)
        sy = np.ascontiguousarray(job["lr_sym"], np.float32)
        keep += [w, sy]
        d.lr_rank = int(job["lr_rank"])
        d.lr_w, d.lr_sym = w.ctypes.data_as(_F), sy.ctypes.data_as(_F)
    d.snaps = int(job.get("snaps", 0))
    d.etas = float(job.get("etas", 0.0))
    d.etar = float(job.get("etar", 0.0))
    return d


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        self.children = weakref.WeakSet()   # open batches and servers
        _chk(lib().cq_open(C.c_int(device), C.byref(self.h)))

    def close(self):
        """Closes the batches and servers still open, then the context."""
        for c in list(self.children):
            c.close()
        if self.h:
            lib().cq_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def plain_forward(self, job):
        keep = []
        d = make_desc(job, keep)
        rec = np.zeros(int(job["nt"]) * d.nr, np.float32)
        _chk(lib().cq_plain_forward(self.h, C.byref(d), rec.ctypes.data_as(_F)))
        return rec.reshape(int(job["nt"]), d.nr)

    def forward_static(self, job):
        """cq_forward_static: the allocation-free variant of plain_forward."""
        keep = []
        d = make_desc(job, keep)
        rec = np.zeros(int(job["nt"]) * d.nr, np.float32)
        _chk(lib().cq_forward_static(self.h, C.byref(d), rec.ctypes.data_as(_F)))
        return rec.reshape(int(job["nt"]), d.nr)

    def batch(self, jobs, tiles=None):
        """A batch of the jobs; tiles: dict of cq_tiles fields (None: the
        engine's choice)."""
        return Batch(self, jobs, tiles)

    def autotune(self, jobs, cache=None):
        """cq_autotune: (tiles dict, ms per step, from_cache)."""
        keep = []
        arr = (JobDesc * len(jobs))()
        for i, j in enumerate(jobs):
            arr[i] = make_desc(j, keep)
        t, ms, hit = Tiles(), C.c_double(), C.c_int()
        _chk(lib().cq_autotune(self.h, arr, C.c_int(len(jobs)), None if cache is None else str(cache).encode(),
                               C.byref(t), C.byref(ms), C.byref(hit)))
        return t.as_dict(), ms.value, bool(hit.value)

    def server(self, max_nxx, max_nzz, hbm_budget=0, max_jobs=0, run_cells=0):
        return Server(self, max_nxx, max_nzz, hbm_budget, max_jobs, run_cells)

    def copy_calibrate(self, nbytes, reps=10):
        """ms per streaming copy of nbytes (reads nbytes, writes nbytes)."""
        ms = C.c_double()
        _chk(lib().cq_copy_calibrate(self.h, C.c_longlong(int(nbytes)), C.c_int(reps), C.byref(ms)))
        return ms.value


class Batch:
    def __init__(self, ctx, jobs, tiles=None):
        self.ctx = ctx
        self.jobs = jobs
        keep = []
        arr = (JobDesc * len(jobs))()
        for i, j in enumerate(jobs):
            arr[i] = make_desc(j, keep)
        self.h = C.c_void_p()
        if tiles is None:
            _chk(lib().cq_batch_create(ctx.h, arr, C.c_int(len(jobs)), C.byref(self.h)))
        else:
            t = Tiles.of(tiles)
            _chk(lib().cq_batch_create_tiled(ctx.h, arr, C.c_int(len(jobs)), C.byref(t), C.byref(self.h)))
        ctx.children.add(self)

    def use_graphs(self, steps_per_graph):
        """cq_batch_use_graphs: run spans of that many steps as HIP graphs (0: off)."""
        _chk(lib().cq_batch_use_graphs(self.h, C.c_int(steps_per_graph)))

    def tiles(self):
        t = Tiles()
        _chk(lib().cq_batch_get_tiles(self.h, C.byref(t)))
        return t.as_dict()

    def step(self, n):
        _chk(lib().cq_batch_step(self.h, C.c_int(n)))

    def sync(self):
        _chk(lib().cq_batch_sync(self.h))

    def run(self):
        _chk(lib().cq_batch_run(self.h))

    def time(self):
        return int(lib().cq_batch_time(self.h))

    def nbytes(self):
        return int(lib().cq_batch_bytes(self.h))

    def rec(self, i):
        j = self.jobs[i]
        nr = len(j["gxg"])
        out = np.zeros(int(j["nt"]) * nr, np.float32)
        _chk(lib().cq_batch_read_rec(self.h, C.c_int(i), out.ctypes.data_as(_F)))
        return out.reshape(int(j["nt"]), nr)

    def p(self, i):
        j = self.jobs[i]
        out = np.zeros(int(j["nxx"]) * int(j["nzz"]), np.float32)
        _chk(lib().cq_batch_read_p(self.h, C.c_int(i), out.ctypes.data_as(_F)))
        return out

    def image(self, i):
        """Raw (img, ssg) of RTM shot i, each nx*nz over the interior."""
        j = self.jobs[i]
        n = (int(j["nxx"]) - 2 * int(j["pml"])) * (int(j["nzz"]) - 2 * int(j["pml"]))
        img = np.zeros(n, np.float32)
        ssg = np.zeros(n, np.float32)
        _chk(lib().cq_batch_read_image(self.h, C.c_int(i), img.ctypes.data_as(_F),
                                       ssg.ctypes.data_as(_F)))
        return img, ssg

    def enable_timing(self, on=True):
        _chk(lib().cq_batch_enable_timing(self.h, C.c_int(1 if on else 0)))

    def pass_times(self):
        r, c, n = C.c_double(), C.c_double(), C.c_int()
        _chk(lib().cq_batch_pass_times(self.h, C.byref(r), C.byref(c), C.byref(n)))
        return r.value, c.value, n.value

    def kernel_times(self):
        """{kernel name: (average ms per launch, launches)} for the passes
        timed while timing was enabled."""
        ms = (C.c_double * NPASS)()
        calls = (C.c_int * NPASS)()
        _chk(lib().cq_batch_kernel_times(self.h, ms, calls))
        out = {}
        for p in range(NPASS):
            if calls[p]:
                out[lib().cq_pass_name(C.c_int(p)).decode()] = (ms[p], calls[p])
        return out

    def close(self):
        if self.h:
            lib().cq_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server (cq_server_* of include/cq.h)."""

    def __init__(self, ctx, max_nxx, max_nzz, hbm_budget=0, max_jobs=0, run_cells=0):
        self.ctx = ctx
        self.jobs = {}
        d = ServerDesc(int(max_nxx), int(max_nzz), int(hbm_budget), int(max_jobs), int(run_cells))
        self.h = C.c_void_p()
        _chk(lib().cq_server_open(ctx.h, C.byref(d), C.byref(self.h)))
        ctx.children.add(self)

    def submit(self, job):
        keep = []
        d = make_desc(job, keep)
        i = C.c_int()
        _chk(lib().cq_server_submit(self.h, C.byref(d), C.byref(i)))
        self.jobs[i.value] = job
        return i.value

    def advance(self, nsteps):
        """Runs up to nsteps global steps; returns (running, queued)."""
        r, q = C.c_int(), C.c_int()
        _chk(lib().cq_server_advance(self.h, C.c_int(int(nsteps)), C.byref(r), C.byref(q)))
        return r.value, q.value

    def state(self, i):
        s = C.c_int()
        _chk(lib().cq_server_poll(self.h, C.c_int(i), C.byref(s)))
        return s.value

    def steps(self, i):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _chk(lib().cq_server_job_steps(self.h, C.c_int(i), C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def result(self, i):
        """Gather (nt, nr) of a modeling job, (img, ssg) of an RTM job."""
        j = self.jobs[i]
        if int(j["kind"]) >= 2:
            n = (int(j["nxx"]) - 2 * int(j["pml"])) * (int(j["nzz"]) - 2 * int(j["pml"]))
            img, ssg = np.zeros(n, np.float32), np.zeros(n, np.float32)
            _chk(lib().cq_server_result(self.h, C.c_int(i), img.ctypes.data_as(_F), ssg.ctypes.data_as(_F)))
            return img, ssg
        nr = len(j["gxg"])
        out = np.zeros(int(j["nt"]) * nr, np.float32)
        _chk(lib().cq_server_result(self.h, C.c_int(i), out.ctypes.data_as(_F), None))
        return out.reshape(int(j["nt"]), nr)

    def release(self, i):
        _chk(lib().cq_server_release(self.h, C.c_int(i)))
        self.jobs.pop(i, None)

    def close(self):
        if self.h:
            lib().cq_server_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
