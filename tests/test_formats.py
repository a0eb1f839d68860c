"""Drop-in formats: the reference's parameter files parse, the geometry
generator writes the same files as the reference's geometry2d.c (compiled
unmodified into oracle/_ref), and the RTM image post-processing follows its
closed forms."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import refc
from cqpy import host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GE = os.path.join(ROOT, "constant-q-pstd_amd", "bin", "ge")
FIXT = os.path.join(ROOT, "tests", "golden")


class Params(C.Structure):
    _fields_ = [("shotnum", C.c_int), ("tmax", C.c_int), ("dt_ms", C.c_float), ("fd", C.c_float),
                ("fo", C.c_float), ("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_float),
                ("dz", C.c_float), ("disx", C.c_float), ("v0", C.c_float), ("etas", C.c_float),
                ("etar", C.c_float), ("fname1", C.c_char * 40), ("fname2", C.c_char * 40),
                ("fname3", C.c_char * 40), ("fname4", C.c_char * 40), ("fname5", C.c_char * 40),
                ("vpname", C.c_char * 80), ("rhoname", C.c_char * 80), ("qname", C.c_char * 80),
                ("recname", C.c_char * 60), ("flagout", C.c_int), ("flagsource", C.c_int),
                ("snapnum", C.c_int), ("pml", C.c_int), ("lr_rank", C.c_int)]


def read_params(path, fmt):
    p = Params()
    rc = host.lib().cqh_read_params(path.encode(), C.c_int(fmt), C.byref(p))
    return rc, p


def test_optional_trailing_keys(tmp_path):
    """Keys appended after the reference's last value: CPML width and the
    low-rank rank; absent, they keep the reference's behaviour."""
    src = open(os.path.join(FIXT, "input", "visac2drealmodeling.txt")).read().rstrip("\n")
    f = tmp_path / "v.txt"
    f.write_text(src + "\n--CPML width (grid points)\n  30\n--low-rank operator rank (0: Taylor form)\n  4\n")
    rc, p = read_params(str(f), 1)
    assert rc == 0 and (p.pml, p.lr_rank) == (30, 4)
    rc, p = read_params(os.path.join(FIXT, "input", "visac2drealmodeling.txt"), 1)
    assert rc == 0 and (p.pml, p.lr_rank) == (10, 0)


def test_reference_parameter_files():
    """The four input files shipped with the reference (copied verbatim into
    tests/golden as data)."""
    rc, p = read_params(os.path.join(FIXT, "input", "visac2drtm.txt"), 3)
    assert rc == 0
    assert (p.shotnum, p.tmax, p.nx, p.nz, p.snapnum, p.flagsource, p.pml) == (80, 2501, 398, 160, 20, 2, 10)
    assert (p.dt_ms, p.fd, p.fo, p.dx, p.dz, p.v0) == (1.0, 30.0, 200.0, 10.0, 10.0, 1500.0)
    assert np.isclose(p.etas, 1e-20) and np.isclose(p.etar, 1e-20)
    assert p.recname == b"../obsdata/recva" and p.qname == b"../output/gasq160_398.bin"
    rc, p = read_params(os.path.join(FIXT, "input", "ac2drealmodeling.txt"), 0)
    assert rc == 0 and p.shotnum == 80 and p.flagout == 0 and p.vpname == b"../output/gasv160_398.bin"
    rc, p = read_params(os.path.join(FIXT, "input", "visac2drealmodeling.txt"), 1)
    assert rc == 0 and p.fo == 200.0 and p.qname == b"../output/gasq160_398.bin"
    rc, p = read_params(os.path.join(FIXT, "input", "ac2drtm.txt"), 2)
    assert rc == 0 and p.v0 == 1500.0 and p.recname == b"../obsdata/recva" and p.snapnum == 20


def test_optional_pml_key(tmp_path):
    src = open(os.path.join(FIXT, "input", "ac2drealmodeling.txt")).read()
    f = tmp_path / "p.txt"
    f.write_text(src.rstrip() + "\n--CPML width (grid points)\n  30\n")
    rc, p = read_params(str(f), 0)
    assert rc == 0 and p.pml == 30 and p.flagsource == 2


def test_missing_file_is_an_error():
    rc, _ = read_params("/nonexistent/x.txt", 0)
    assert rc != 0


def run_ge(exe, workdir, params):
    os.makedirs(workdir, exist_ok=True)
    shutil.copy(params, os.path.join(workdir, "parameters.txt"))
    subprocess.run([exe], cwd=workdir, check=True, stdout=subprocess.DEVNULL)
    return {f: open(os.path.join(workdir, f), "rb").read() for f in sorted(os.listdir(workdir))
            if f.endswith(".bin")}


def test_geometry_generator_matches_reference(tmp_path):
    ref = os.path.join(ROOT, "oracle", "_ref", "ge_ref")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref/ge_ref not built (reference tree absent)")
    par = os.path.join(FIXT, "geometry_parameters.txt")
    a = run_ge(GE, str(tmp_path / "mine"), par)
    b = run_ge(ref, str(tmp_path / "ref"), par)
    assert a.keys() == b.keys() and len(a) == 2 * 80 + 3
    for k in a:
        assert a[k] == b[k], k
    tn = np.frombuffer(a["tracenum.bin"], np.int32)
    assert tn[0] == 191   # shot 0 at x = 400 m, offset 1500 m, 10 m traces


def test_filter_and_mute_closed_form():
    nx, nz, dx, dz = 9, 7, 10.0, 5.0
    x = np.arange(nx)[:, None] * dx
    z = np.arange(nz)[None, :] * dz
    img = (3 * x ** 2 + 2 * z ** 2).astype(np.float32)   # Laplacian = 6 + 4
    out = np.zeros(nx * nz, np.float32)
    F = C.POINTER(C.c_float)
    host.lib().cqh_filter_image(C.c_int(nx), C.c_int(nz), C.c_float(dx), C.c_float(dz),
                                np.ascontiguousarray(img).ctypes.data_as(F), out.ctypes.data_as(F))
    out = out.reshape(nx, nz)
    assert np.allclose(out[1:-1, 1:-1], 10.0, rtol=1e-4)
    assert not out[0].any() and not out[-1].any() and not out[:, 0].any() and not out[:, -1].any()
    m = np.ones(nx * nz, np.float32)
    host.lib().cqh_mute_image(C.c_int(nx), C.c_int(nz), C.c_int(3), C.c_int(5), C.c_int(9),
                              m.ctypes.data_as(F))
    m = m.reshape(nx, nz)
    assert not m[:2].any() and m[2:6].all() and not m[6:].any()
