"""The drop-in drivers end to end on a small generated dataset: reference-format
parameter files, geometry files and float32 models in, reference-named
outputs out.  aps / vaps gathers match the FP64 oracle on the same shots
(rel-L2 <= 1e-4), snapshots (flagout = 1) have the reference's layout and
equal the engine's pressure field, and artm / qrtm write [image |
illumination] stacks equal to the ones assembled from the FP64 oracle.  The
reference's own shipped parameter files and geometry run unchanged but for
the time axis (test_shipped_default_run)."""
import os
import subprocess

import numpy as np
import pytest

import orc
from cqpy import engine, host, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "constant-q-pstd_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden", "input")
NXT, NZT, DX, NT, DT_MS, FD, FO, PML = 64, 32, 10.0, 100, 1.0, 30.0, 60.0, 10
SHOTS = [20, 32, 44]
HALF = 15                    # receivers at s-15 .. s+15 (cells)


def param_file(path, golden, values):
    labels = [l.rstrip("\n") for l in open(os.path.join(GOLD, golden))][0::2]
    assert len(labels) == len(values), (golden, len(labels), len(values))
    with open(path, "w") as f:
        for lab, v in zip(labels, values):
            f.write("%s\n  %s\n" % (lab, v))
        f.write("--CPML width (grid points)\n  %d\n" % PML)


@pytest.fixture(scope="module")
def tree(tmp_path_factory):
    t = tmp_path_factory.mktemp("cq")
    for d in ("work", "input", "geometry", "output", "obsdata"):
        (t / d).mkdir()
    vp, rho, q = synth.layered_model(NXT, NZT, DX, seed=77, nlayers=3, vmin=1500.0, vmax=3000.0,
                                     qrange=(30.0, 120.0))
    vp.tofile(t / "output" / "vp.bin")
    rho.tofile(t / "output" / "rho.bin")
    q.tofile(t / "output" / "q.bin")
    np.full(len(SHOTS), 2 * HALF + 1, np.int32).tofile(t / "geometry" / "tracenum.bin")
    np.array([s * DX for s in SHOTS], np.float32).tofile(t / "geometry" / "shotxp.bin")
    np.full(len(SHOTS), 2 * DX, np.float32).tofile(t / "geometry" / "shotzp.bin")
    for i, s in enumerate(SHOTS):
        (np.arange(s - HALF, s + HALF + 1) * DX).astype(np.float32).tofile(t / "geometry" / ("gxp%d.bin" % i))
        np.full(2 * HALF + 1, 2 * DX, np.float32).tofile(t / "geometry" / ("gzp%d.bin" % i))
    geo = ["../geometry/tracenum.bin", "../geometry/shotxp.bin", "../geometry/shotzp.bin",
           "../geometry/gxp", "../geometry/gzp"]
    head = [len(SHOTS), NT, DT_MS, FD]
    grid = [NXT, NZT, DX, DX, 0.0]
    param_file(t / "input" / "aps.txt", "ac2drealmodeling.txt",
               head + grid + geo + ["../output/vp.bin", "../output/rho.bin", 1, 2])
    param_file(t / "input" / "vaps.txt", "visac2drealmodeling.txt",
               head + [FO] + grid + geo + ["../output/vp.bin", "../output/rho.bin", "../output/q.bin", 1, 2])
    param_file(t / "input" / "vaps_lr.txt", "visac2drealmodeling.txt",
               head + [FO] + grid + geo + ["../output/vp.bin", "../output/rho.bin", "../output/q.bin", 0, 2])
    with open(t / "input" / "vaps_lr.txt", "a") as f:
        f.write("--low-rank operator rank (0: Taylor form)\n  4\n")
    param_file(t / "input" / "artm.txt", "ac2drtm.txt",
               head + grid + [1500.0] + geo + ["../output/vp.bin", "../output/rho.bin", "../obsdata/reca", 2, 12])
    param_file(t / "input" / "qrtm.txt", "visac2drtm.txt",
               head + [FO] + grid + [1500.0, 1e-20, 1e-20] + geo +
               ["../output/vp.bin", "../output/rho.bin", "../output/q.bin", "../obsdata/recva", 2, 12])
    return t, vp, rho, q


def run(t, mode, batch=2):
    print("driver", mode, flush=True)
    r = subprocess.run([os.path.join(BIN, mode), "../input/%s.txt" % mode, "../obsdata", str(batch)],
                       cwd=t / "work", capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def shot_job(kind, vp, rho, q, s):
    gxp = (np.arange(s - HALF, s + HALF + 1) * DX).astype(np.float32)
    gzp = np.full(2 * HALF + 1, 2 * DX, np.float32)
    return synth.make_job(kind, vp, rho, NXT, NZT, DX, DX, DT_MS / 1000.0, NT, FD, s * DX, 2 * DX, gxp, gzp,
                          PML, qt=q, fo=FO)


@pytest.mark.parametrize("mode,kind,prefix", [("aps", 0, "reca"), ("vaps", 1, "recva")])
def test_modeling_driver_vs_oracle(tree, mode, kind, prefix):
    t, vp, rho, q = tree
    run(t, mode)
    nr = 2 * HALF + 1
    for i, s in enumerate(SHOTS):
        got = np.fromfile(t / "obsdata" / ("%s%d_%d.bin" % (prefix, i, nr)), np.float32).reshape(NT, nr)
        job = shot_job(kind, vp, rho, q, s)
        e = engine.rel_l2(got, orc.forward(job))
        print(mode, "shot", i, "rel-L2 vs FP64 oracle", e)
        assert e <= 1e-4
    # snapshots: interior of p after every 10th step, x-major, appended
    job = shot_job(kind, vp, rho, q, SHOTS[0])
    nx, nz = job["nxx"] - 2 * PML, job["nzz"] - 2 * PML
    snap = np.fromfile(t / "obsdata" / ("%s%d_%d_%d.bin" % ("snapa" if kind == 0 else "snapva", 0, nz, nx)),
                       np.float32)
    assert snap.size == (NT // 10) * nx * nz
    with engine.Context(0) as ctx:
        b = ctx.batch([job])
        b.step(NT - 9)   # through t = 90
        p = b.p(0).reshape(job["nxx"], job["nzz"])[PML:-PML, PML:-PML]
        b.close()
    last = snap.reshape(NT // 10, nx, nz)[-1]
    e = engine.rel_l2(last, p)
    print(mode, "last snapshot vs engine p", e)
    assert e <= 1e-5


def expected_stack(t, kind, vp, rho, q, snaps, v0, eta):
    """The [image | illumination] stack the reference's RTM driver builds
    (sgpstd-qrtm/acps4.cu:909-945, main.c:470-480), assembled here from the
    FP64 oracle's raw image of each shot: observed gather read back from the
    modeling run, direct wave removed, Laplacian filter, mute outside the
    receiver span, projection to the model, sum over shots."""
    img_sp = np.zeros(NXT * NZT, np.float32)
    ssg_sp = np.zeros(NXT * NZT, np.float32)
    nr = 2 * HALF + 1
    for i, s in enumerate(SHOTS):
        job = shot_job(kind, vp, rho, q, s)
        gxp = (np.arange(s - HALF, s + HALF + 1) * DX).astype(np.float32)
        gzp = np.full(nr, 2 * DX, np.float32)
        obs = np.fromfile(t / "obsdata" / ("%s%d_%d.bin" % ("reca" if kind == 2 else "recva", i, nr)), np.float32)
        obs = host.removedirect(obs, NT, nr, DT_MS / 1000.0, FD, s * DX, 2 * DX, gxp, gzp, v0).reshape(NT, nr)
        job.update(obs=obs, snaps=snaps, etas=eta, etar=eta)
        img, ssg = orc.rtm(job, host.revolve_trace(NT, snaps)[0])
        nx, nz = job["nxx"] - 2 * PML, job["nzz"] - 2 * PML
        imgf = host.filter_image(img.astype(np.float32).ravel(), nx, nz, DX, DX)
        imgf = host.mute_image(imgf, nx, nz, PML, job["lgest"], job["rgest"])
        # projectomodel writes the window into a zeroed model, then the stack adds it
        img_sp += host.projectomodel(imgf, NXT, NZT, nx, nz, PML, job["mlx"], DX)
        ssg_sp += host.projectomodel(ssg.astype(np.float32).ravel(), NXT, NZT, nx, nz, PML, job["mlx"], DX)
    return img_sp, ssg_sp


def test_vaps_low_rank_key(tree):
    """The optional trailing key "low-rank 4" makes vaps decompose each shot's
    operator (fractional-Laplacian form, seed 1 + shot id) and apply it at
    rank 4: gathers match the oracle's low-rank apply of the same
    decomposition (rel-L2 <= 1e-4)."""
    t, vp, rho, q = tree
    out = t / "obsdata_lr"
    out.mkdir(exist_ok=True)
    print("driver vaps low-rank", flush=True)
    r = subprocess.run([os.path.join(BIN, "vaps"), "../input/vaps_lr.txt", str(out), "2"], cwd=t / "work",
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    nr = 2 * HALF + 1
    for i, s in enumerate(SHOTS):
        got = np.fromfile(out / ("recva%d_%d.bin" % (i, nr)), np.float32).reshape(NT, nr)
        job = shot_job(1, vp, rho, q, s)
        lr = host.lowrank(job, 4, symbol=host.LR_FRACLAP, seed=1 + i)
        e = engine.rel_l2(got, orc.forward(job, lowrank=lr))
        print("vaps low-rank shot", i, "rel-L2 vs FP64 oracle", e)
        assert e <= 1e-4


@pytest.mark.parametrize("mode,out,kind", [("artm", "RTM.BIN", 2), ("qrtm", "QRTM.BIN", 3)])
def test_rtm_driver_stack_vs_oracle(tree, mode, out, kind):
    """artm / qrtm write [image | illumination] of the model size; both equal
    the stack assembled from the FP64 oracle (image rel-L2 <= 1e-3,
    illumination <= 1e-4).  The driver runs the three shots two at a time."""
    t, vp, rho, q = tree
    run(t, "aps" if mode == "artm" else "vaps")   # observed data
    run(t, mode)
    got = np.fromfile(t / "work" / out, np.float32)
    assert got.size == 2 * NXT * NZT
    img, ssg = got[:NXT * NZT], got[NXT * NZT:]
    assert np.isfinite(img).all() and np.isfinite(ssg).all()
    ref_img, ref_ssg = expected_stack(t, kind, vp, rho, q, 12, 1500.0, 1e-20)
    ei, es = engine.rel_l2(img, ref_img), engine.rel_l2(ssg, ref_ssg)
    print(mode, "stack rel-L2 vs oracle: image", ei, "illumination", es)
    assert es <= 1e-4
    assert ei <= 1e-3


# ---------------------------------------------------------------------------
# the reference's shipped inputs, unchanged but for the time axis
# ---------------------------------------------------------------------------
SHIP_NT = 500


@pytest.fixture(scope="module")
def shipped(tmp_path_factory):
    """The reference's directory layout with its shipped parameter files
    (tests/golden/input, only the time sample number cut from 2501 to
    SHIP_NT) and its geometry generator's parameters
    (tests/golden/geometry_parameters.txt: 80 shots from 400 m every 40 m,
    1500 m offset), run through bin/ge; the 398 x 160 models are synthetic
    (seeded layered vp up to 3500 m/s, Gardner density, Q 30-200)."""
    t = tmp_path_factory.mktemp("ship")
    for d in ("work", "input", "geometry", "output", "obsdata"):
        (t / d).mkdir()
    (t / "geometry" / "parameters.txt").write_text(
        open(os.path.join(ROOT, "tests", "golden", "geometry_parameters.txt")).read())
    r = subprocess.run([os.path.join(BIN, "ge"), "parameters.txt"], cwd=t / "geometry", capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    vp, rho, q = synth.layered_model(398, 160, 10.0, seed=61, vmin=1500.0, vmax=3500.0, qrange=(30.0, 200.0))
    vp.tofile(t / "output" / "gasv160_398.bin")
    rho.tofile(t / "output" / "rho160_398.bin")
    q.tofile(t / "output" / "gasq160_398.bin")
    for name in ("ac2drealmodeling.txt", "visac2drealmodeling.txt"):
        lines = open(os.path.join(GOLD, name)).read().split("\n")
        k = [i for i, l in enumerate(lines) if l.startswith("--time sample number")][0]
        lines[k + 1] = "  %d" % SHIP_NT
        (t / "input" / name).write_text("\n".join(lines))
    return t, vp, rho, q


@pytest.mark.parametrize("mode,kind,prefix,shots", [("aps", 0, "reca", (0, 79)), ("vaps", 1, "recva", (40,))])
def test_shipped_default_run(shipped, mode, kind, prefix, shots):
    """bin/aps and bin/vaps with no arguments read ../input/<reference
    name>.txt, size the reference's window (301 x 160, padded 321 x 180 with
    its pml = 10; 321 = 3 107 runs by Bluestein), run all 80 shots and write
    ../obsdata/<reca|recva><shot>_<nr>.bin; sampled shots match the FP64
    oracle built from the same geometry files (rel-L2 <= 1e-4)."""
    t, vp, rho, q = shipped
    print("driver", mode, "default input", flush=True)
    r = subprocess.run([os.path.join(BIN, mode)], cwd=t / "work", capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "window 301 x 160 (padded 321 x 180), 80 shots" in r.stdout, r.stdout[:400]
    tracenum = np.fromfile(t / "geometry" / "tracenum.bin", np.int32)
    shotxp = np.fromfile(t / "geometry" / "shotxp.bin", np.float32)
    shotzp = np.fromfile(t / "geometry" / "shotzp.bin", np.float32)
    assert tracenum.size == 80
    for s in shots:
        nr = int(tracenum[s])
        got = np.fromfile(t / "obsdata" / ("%s%d_%d.bin" % (prefix, s, nr)), np.float32).reshape(SHIP_NT, nr)
        gxp = np.fromfile(t / "geometry" / ("gxp%d.bin" % s), np.float32)
        gzp = np.fromfile(t / "geometry" / ("gzp%d.bin" % s), np.float32)
        job = synth.make_job(kind, vp, rho, 398, 160, 10.0, 10.0, 0.001, SHIP_NT, 30.0, float(shotxp[s]),
                             float(shotzp[s]), gxp, gzp, 10, qt=q, fo=200.0, sroffmx=1500.0)
        assert (job["nxx"], job["nzz"]) == (321, 180)
        e = engine.rel_l2(got, orc.forward(job))
        print(mode, "shot", s, "nr", nr, "rel-L2 vs FP64 oracle", e, flush=True)
        assert e <= 1e-4


@pytest.mark.parametrize("mode", ["aps", "vaps", "artm", "qrtm"])
def test_shims_equal_drivers(tree, tmp_path, mode):
    """The four acforward bodies of libcq_shim.so (cq_shim_aps / vaps / artm /
    qrtm; sgpstd-qmodeling/acps4.cu:184-198, acsglw.cu:185-200,
    sgpstd-qrtm/acps4.cu:183-198, acsglw.cu:189-204) on the three shots of the
    generated dataset: the modeling shims write gather files equal, bit for
    bit, to bin/aps's / bin/vaps's; the RTM shims add to img_sp / ssg_sp
    exactly the stack bin/artm / bin/qrtm write (shots summed in the same
    order, every shot's image independent of its batch)."""
    import ctypes as C
    t, vp, rho, q = tree
    q_kind = mode in ("vaps", "qrtm")
    rtm = mode in ("artm", "qrtm")
    if rtm:
        run(t, "vaps" if q_kind else "aps")   # observed data
    run(t, mode)
    n = len(SHOTS)
    tracenum = np.fromfile(t / "geometry" / "tracenum.bin", np.int32)
    sx = np.fromfile(t / "geometry" / "shotxp.bin", np.float32)
    sz = np.fromfile(t / "geometry" / "shotzp.bin", np.float32)
    j = shot_job(1 if q_kind else 0, vp, rho, q, SHOTS[0])
    bbx, bbz = host.bbx_bbz(NXT, NZT, np.ascontiguousarray(rho, np.float32).ravel())
    keep = []

    def F(a):
        keep.append(np.ascontiguousarray(a, np.float32))
        return keep[-1].ctypes.data_as(C.POINTER(C.c_float))

    def I(a):
        keep.append(np.ascontiguousarray(a, np.int32))
        return keep[-1].ctypes.data_as(C.POINTER(C.c_int))

    head = [C.c_int(NT), C.c_float(DT_MS / 1000.0), F(j["rik"]), C.c_int(n), I(np.arange(n)), F(sx), F(sz),
            C.c_int(j["nx"]), C.c_int(j["nz"]), C.c_int(j["nxx"]), C.c_int(j["nzz"]), C.c_int(PML),
            C.c_float(DX), C.c_float(DX), C.c_float(j["sroffmx"]), C.c_float(0.0), F(vp)]
    if q_kind:
        tq = host.transfermodel(FD, FO, np.ascontiguousarray(vp, np.float32).ravel(),
                                np.ascontiguousarray(q, np.float32).ravel())
        head += [F(tq["vwd"]), F(tq["tao"]), F(tq["lam"]), F(tq["gam"])]
    head += [F(rho), F(bbx), F(bbz), C.c_int(NXT), C.c_int(NZT), I(tracenum)]
    coef = [F(j[k]) for k in ("iax", "ibx", "iaz", "ibz", "hax", "hbx", "haz", "hbz",
                              "kxsmr", "kxscr", "kzsmr", "kzscr", "kxsmi", "kxsci", "kzsmi", "kzsci")]
    names = [str(t / "geometry" / "gxp").encode(), str(t / "geometry" / "gzp").encode()]
    lib = C.CDLL(os.path.join(ROOT, "constant-q-pstd_amd", "libcq_shim.so"))
    print("case shim", mode, flush=True)
    if not rtm:
        rc = getattr(lib, "cq_shim_" + mode)(*head, C.c_int(0), *coef, *names, str(tmp_path).encode())
        assert rc == 0, rc
        prefix = "recva" if q_kind else "reca"
        for i in range(n):
            name = "%s%d_%d.bin" % (prefix, i, int(tracenum[i]))
            assert (tmp_path / name).read_bytes() == (t / "obsdata" / name).read_bytes(), name
        return
    img = np.zeros(NXT * NZT, np.float32)
    ssg = np.zeros(NXT * NZT, np.float32)
    rtm_tail = [img.ctypes.data_as(C.POINTER(C.c_float)), ssg.ctypes.data_as(C.POINTER(C.c_float)),
                C.c_float(FD), C.c_float(1500.0)]
    if q_kind:
        rtm_tail += [C.c_float(1e-20), C.c_float(1e-20)]
    recname = str(t / "obsdata" / ("recva" if q_kind else "reca")).encode()
    rc = getattr(lib, "cq_shim_" + mode)(C.c_int(12), *head, *coef, *rtm_tail, *names, recname)
    assert rc == 0, rc
    got = np.fromfile(t / "work" / ("QRTM.BIN" if q_kind else "RTM.BIN"), np.float32)
    assert np.abs(img).max() > 0
    assert np.array_equal(got[:NXT * NZT], img) and np.array_equal(got[NXT * NZT:], ssg)


def test_shim_aps_equals_driver(shipped, tmp_path):
    """cq_shim_aps (libcq_shim.so, the body of the reference's aps acforward,
    sgpstd-qmodeling/acps4.cu:184-750, for shots given as arrays) on three
    shots of the shipped run writes reca<shot>_<nr>.bin files equal, bit for
    bit, to those of bin/aps (same window, same tiles), and snapshots with
    flagout = 1 (50 frames of the 301 x 160 interior for 500 steps)."""
    import ctypes as C
    t, vp, rho, q = shipped
    shots = [0, 40, 79]
    tracenum = np.fromfile(t / "geometry" / "tracenum.bin", np.int32)
    shotxp = np.fromfile(t / "geometry" / "shotxp.bin", np.float32)
    shotzp = np.fromfile(t / "geometry" / "shotzp.bin", np.float32)
    if not all((t / "obsdata" / ("reca%d_%d.bin" % (s, int(tracenum[s])))).exists() for s in shots):
        r = subprocess.run([os.path.join(BIN, "aps")], cwd=t / "work", capture_output=True, text=True, timeout=280)
        assert r.returncode == 0, r.stdout + r.stderr
    gxp = np.fromfile(t / "geometry" / "gxp0.bin", np.float32)
    gzp = np.fromfile(t / "geometry" / "gzp0.bin", np.float32)
    j = synth.make_job(0, vp, rho, 398, 160, 10.0, 10.0, 0.001, SHIP_NT, 30.0, float(shotxp[0]), float(shotzp[0]),
                       gxp, gzp, 10, sroffmx=1500.0, want_k2=False)
    bbx, bbz = host.bbx_bbz(398, 160, np.ascontiguousarray(rho, np.float32).ravel())
    lib = C.CDLL(os.path.join(ROOT, "constant-q-pstd_amd", "libcq_shim.so"))
    keep = {}

    def arr(name, a, dt=np.float32):
        keep[name] = np.ascontiguousarray(a, dt)
        ct = C.c_float if dt == np.float32 else C.c_int
        return keep[name].ctypes.data_as(C.POINTER(ct))

    for flagout, out in ((0, tmp_path / "o0"), (1, tmp_path / "o1")):
        out.mkdir()
        print("case shim flagout", flagout, flush=True)
        rc = lib.cq_shim_aps(
            C.c_int(SHIP_NT), C.c_float(0.001), arr("rik", j["rik"]), C.c_int(len(shots)),
            arr("ids", shots, np.int32), arr("sx", shotxp[shots]), arr("sz", shotzp[shots]),
            C.c_int(j["nx"]), C.c_int(j["nz"]), C.c_int(j["nxx"]), C.c_int(j["nzz"]), C.c_int(10),
            C.c_float(10.0), C.c_float(10.0), C.c_float(1500.0), C.c_float(0.0),
            arr("vpt", vp), arr("rhot", rho), arr("bbxt", bbx), arr("bbzt", bbz), C.c_int(398), C.c_int(160),
            arr("tn", tracenum, np.int32), C.c_int(flagout),
            *[arr(k, j[k]) for k in ("iax", "ibx", "iaz", "ibz", "hax", "hbx", "haz", "hbz",
                                     "kxsmr", "kxscr", "kzsmr", "kzscr", "kxsmi", "kxsci", "kzsmi", "kzsci")],
            str(t / "geometry" / "gxp").encode(), str(t / "geometry" / "gzp").encode(), str(out).encode())
        assert rc == 0, rc
        for s in shots:
            name = "reca%d_%d.bin" % (s, int(tracenum[s]))
            assert (out / name).read_bytes() == (t / "obsdata" / name).read_bytes(), name
        if flagout:
            snap = np.fromfile(out / ("snapa%d_%d_%d.bin" % (shots[1], j["nz"], j["nx"])), np.float32)
            assert snap.size == (SHIP_NT + 9) // 10 * j["nx"] * j["nz"]
            assert np.isfinite(snap).all() and np.abs(snap).max() > 0


@pytest.mark.parametrize("split,total", [("1,2", 2), ("2,1", 3)])
def test_reference_shot_split(tree, split, total):
    """CQ_SHOT_SPLIT="numprocs,GPU_N" runs the shots the reference's MPI split
    would run (acousticsgps4.c:255-273): 3 shots on 1 rank x 2 GPUs run shots
    0 and 1 only (the remainder shot is dropped), on 2 ranks x 1 GPU all three.
    Every shot that runs writes the gather of the default run."""
    t, vp, rho, q = tree
    out = t / ("obsdata_split_" + split.replace(",", "_"))
    out.mkdir(exist_ok=True)
    print("driver aps split", split, flush=True)
    r = subprocess.run([os.path.join(BIN, "aps"), "../input/aps.txt", str(out), "2"], cwd=t / "work",
                       capture_output=True, text=True, timeout=240, env=dict(os.environ, CQ_SHOT_SPLIT=split))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Run total shot number is %d" % total in r.stdout
    nr = 2 * HALF + 1
    for i in range(len(SHOTS)):
        f = out / ("reca%d_%d.bin" % (i, nr))
        if i >= total:
            assert not f.exists()
            continue
        got = np.fromfile(f, np.float32).reshape(NT, nr)
        e = engine.rel_l2(got, orc.forward(shot_job(0, vp, rho, q, SHOTS[i])))
        print("split", split, "shot", i, "rel-L2 vs FP64 oracle", e)
        assert e <= 1e-4
