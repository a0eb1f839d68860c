"""Host coefficient builders and windowing vs the reference's own host C
compiled unmodified into oracle/_ref (bit-exact, since the engine is fed these
float32 arrays exactly as the reference's kernels are)."""
import ctypes as C

import numpy as np
import pytest

import refc
from cqpy import host

F = C.POINTER(C.c_float)


def fp(a):
    return a.ctypes.data_as(F)


GRIDS = [(459, 220, 30, 10.0, 10.0, 0.0005, 15.0), (321, 180, 10, 10.0, 10.0, 0.001, 30.0),
         (64, 48, 8, 5.0, 7.5, 0.0004, 20.0), (1281, 680, 40, 5.0, 5.0, 0.0004, 20.0)]


@pytest.mark.parametrize("nxx,nzz,pml,dx,dz,dt,fd", GRIDS)
def test_pml_coeffs_bitexact(nxx, nzz, pml, dx, dz, dt, fd):
    ref = refc.load("libref_qm_funca.so")
    mine = host.pml_coeffs(pml, fd, nxx, nzz, dx, dz, dt)
    r = {k: np.zeros(nxx, np.float32) for k in ("iax", "ibx", "hax", "hbx")}
    r.update({k: np.zeros(nzz, np.float32) for k in ("iaz", "ibz", "haz", "hbz")})
    ref.get_pml_parameters(C.c_int(pml), C.c_float(fd), fp(r["iax"]), fp(r["ibx"]), fp(r["iaz"]),
                           fp(r["ibz"]), fp(r["hax"]), fp(r["hbx"]), fp(r["haz"]), fp(r["hbz"]),
                           C.c_int(nxx), C.c_int(nzz), C.c_float(dx), C.c_float(dz),
                           C.c_float(dt))
    for k in r:
        assert np.array_equal(mine[k].view(np.uint32), r[k].view(np.uint32)), k


@pytest.mark.parametrize("nxx,nzz,pml,dx,dz,dt,fd", GRIDS)
def test_shift_coeffs_bitexact(nxx, nzz, pml, dx, dz, dt, fd):
    ref = refc.load("libref_qm_funca.so")
    mine = host.shift_coeffs(dx, dz, nxx, nzz)
    r = {k: np.zeros(nxx, np.float32) for k in ("kxsmr", "kxscr", "kxsmi", "kxsci")}
    r.update({k: np.zeros(nzz, np.float32) for k in ("kzsmr", "kzscr", "kzsmi", "kzsci")})
    ref.cal_shift_parameters(fp(r["kxsmr"]), fp(r["kxscr"]), fp(r["kzsmr"]), fp(r["kzscr"]),
                             fp(r["kxsmi"]), fp(r["kxsci"]), fp(r["kzsmi"]), fp(r["kzsci"]),
                             C.c_float(dx), C.c_float(dz), C.c_int(nxx), C.c_int(nzz))
    for k in r:
        assert np.array_equal(mine[k].view(np.uint32), r[k].view(np.uint32)), k


@pytest.mark.parametrize("nxx,nzz", [(459, 220), (65, 33), (128, 96)])
def test_k2d_bitexact(nxx, nzz):
    refa = refc.load("libref_qm_funca.so")
    refq = refc.load("libref_qm_func.so")
    mine = host.k2d(10.0, 7.5, nxx, nzz, q=True)
    k1 = np.zeros(nxx * nzz, np.float32)
    refa.cal_k2d(C.c_float(10.0), C.c_float(7.5), C.c_int(nxx), C.c_int(nzz), fp(k1))
    assert np.array_equal(mine["k2d1"].view(np.uint32), k1.view(np.uint32))
    r = {k: np.zeros(nxx * nzz, np.float32) for k in ("k2d1", "k2d2", "k2d3", "k2da")}
    refq.cal_k2d(C.c_float(10.0), C.c_float(7.5), C.c_int(nxx), C.c_int(nzz), fp(r["k2d1"]),
                 fp(r["k2d2"]), fp(r["k2d3"]), fp(r["k2da"]))
    for k in r:
        assert np.array_equal(mine[k].view(np.uint32), r[k].view(np.uint32)), k


def test_bbx_ricker_transfer_bitexact():
    ref = refc.load("libref_qm_funca.so")
    refq = refc.load("libref_qm_func.so")
    rng = np.random.default_rng(5)
    nxx, nzz = 37, 29
    rho = rng.uniform(1000, 2600, nxx * nzz).astype(np.float32)
    bx, bz = host.bbx_bbz(nxx, nzz, rho)
    rbx, rbz = np.zeros_like(rho), np.zeros_like(rho)
    ref.cal_bbx_bbz(C.c_int(nxx), C.c_int(nzz), fp(rho), fp(rbx), fp(rbz))
    assert np.array_equal(bx, rbx) and np.array_equal(bz, rbz)
    for flag in (2, 1):
        mine = host.ricker(3000, 0.0005, 15.0, flag)
        r = np.zeros(3000, np.float32)
        ref.rik_cal(C.c_int(3000), C.c_float(0.0005), C.c_float(15.0), fp(r), C.c_int(flag))
        assert np.array_equal(mine.view(np.uint32), r.view(np.uint32))
    vp = rng.uniform(1500, 4500, 500).astype(np.float32)
    qp = rng.uniform(15, 300, 500).astype(np.float32)
    mine = host.transfermodel(15.0, 200.0, vp, qp)
    r = {k: np.zeros(500, np.float32) for k in ("gam", "lam", "vwd", "tao")}
    refq.transfermodel(C.c_int(500), C.c_int(1), C.c_float(15.0), C.c_float(200.0), fp(vp),
                       fp(qp), fp(r["gam"]), fp(r["lam"]), fp(r["vwd"]), fp(r["tao"]))
    for k in r:
        assert np.array_equal(mine[k].view(np.uint32), r[k].view(np.uint32)), k
    refq.get_max_tao.restype = C.c_float
    assert host.get_max_tao(r["tao"]) == refq.get_max_tao(C.c_int(1), C.c_int(500), fp(r["tao"]))
    ref.check_stability_ps.restype = C.c_float
    assert host.check_stability(4000.0, 0.0005, 10.0, 10.0) == ref.check_stability_ps(
        C.c_float(4000.0), C.c_float(0.0005), C.c_float(10.0), C.c_float(10.0))


@pytest.mark.parametrize("mlx_cells,nx", [(0.0, 40), (3.0, 40), (-5.0, 40), (30.0, 40),
                                          (-12.0, 80), (2.4, 50), (-2.6, 30), (7.0, 20)])
def test_window_bitexact(mlx_cells, nx):
    ref = refc.load("libref_qm_funca.so")
    rng = np.random.default_rng(11)
    nxt, nzt, pml, dx = 50, 23, 6, 10.0
    full = rng.standard_normal(nxt * nzt).astype(np.float32)
    mlx = mlx_cells * dx
    mine = host.window_field(full, nxt, nzt, nx, nzt, pml, mlx, dx)
    nxx, nzz = nx + 2 * pml, nzt + 2 * pml
    r = np.zeros(nxx * nzz, np.float32)
    ref.select_part_model(C.c_int(pml), C.c_float(mlx), C.c_float(dx), C.c_int(nxt),
                          C.c_int(nzt), C.c_int(nx), C.c_int(nzt), C.c_int(nxx), C.c_int(nzz),
                          fp(full), fp(r))
    ref.expand_pml_areas(C.c_int(nx), C.c_int(nzt), C.c_int(nxx), C.c_int(nzz), C.c_int(pml),
                         fp(r))
    assert np.array_equal(mine, r)
    interior = rng.standard_normal(nx * nzt).astype(np.float32)
    pm = host.projectomodel(interior, nxt, nzt, nx, nzt, pml, mlx, dx)
    pr = np.zeros(nxt * nzt, np.float32)
    ref.projectomodel(C.c_int(pml), C.c_float(mlx), C.c_float(dx), C.c_int(nxt), C.c_int(nzt),
                      C.c_int(nx), C.c_int(nzt), C.c_int(nxx), C.c_int(nzz), fp(pr),
                      fp(interior))
    assert np.array_equal(pm, pr)


def test_removedirect_bitexact():
    ref = refc.load("libref_qr_funca.so")
    rng = np.random.default_rng(3)
    nr, tmax = 57, 900
    obs = rng.standard_normal(tmax * nr).astype(np.float32)
    gxp = np.linspace(0, 2000, nr).astype(np.float32)
    gzp = np.full(nr, 20.0, np.float32)
    mine = host.removedirect(obs, tmax, nr, 0.001, 25.0, 990.0, 10.0, gxp, gzp, 1500.0)
    r = obs.copy()
    ref.removedirect(C.c_int(nr), C.c_int(tmax), C.c_int(nr), C.c_float(0.001),
                     C.c_float(25.0), C.c_float(990.0), C.c_float(10.0), fp(gxp), fp(gzp),
                     C.c_float(1500.0), fp(r))
    assert np.array_equal(mine, r)
