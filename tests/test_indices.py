"""Known-answer cases of the source / receiver grid indexing
(cqh_shot_indices, restating sgpstd-qmodeling/acps4.cu:391-426), worked out
by hand from that code's integer rules: xsg = (int)((sroffmx + disx)/dx +
0.5) + pml, zsg = (int)(zs/dz + 0.5) + pml, gxg = xsg +/- (int)(|gx - xs|/dx +
0.5), receivers clamped into [pml, nxx - pml - 1] (counted), lgest / rgest
the extreme receiver columns, and an error for a source outside the window.
Integer results, checked exactly; the expected values below are literals,
not recomputed."""
import numpy as np
import pytest

from cqpy import host


def idx(pml, nxx, nzz, sroffmx, disx, xs, zs, gx, gz, dx=10.0, dz=10.0):
    return host.shot_indices(pml, nxx, nzz, dx, dz, sroffmx, disx, xs, zs,
                             np.array(gx, np.float32), np.array(gz, np.float32))


def test_rounding_half_up_both_sides():
    # window 2*1500/10 + 1 = 301 columns + 2*10 CPML = 321 (the reference's default)
    gx = [2000.0, 2004.9, 2005.0, 1995.1, 1994.9, 2014.9, 1985.0]
    gz = [0.0, 20.0, 4.9, 5.0, 14.9, 15.0, 0.0]
    gxg, gzg, g = idx(10, 321, 180, 1500.0, 0.0, 2000.0, 0.0, gx, gz)
    assert (g.xsg, g.zsg) == (160, 10)
    # +0.49 -> 0, +0.5 -> 1, -0.49 -> 0, -0.51 -> -1, +1.49 -> 1, -1.5 -> -2
    assert gxg.tolist() == [160, 160, 161, 160, 159, 161, 158]
    assert gzg.tolist() == [10, 12, 10, 11, 11, 12, 10]
    assert (g.lgest, g.rgest, g.nclamped) == (158, 161, 0)
    assert g.mlx == 500.0


def test_clamping_into_the_window_interior():
    gx = [500.0, 3500.0, 400.0, 3600.0, 300.0]
    gxg, gzg, g = idx(10, 321, 180, 1500.0, 0.0, 2000.0, 20.0, gx, [0.0] * 5)
    assert g.zsg == 12
    # 500 and 3500 land exactly on the first / last interior column; 400, 300
    # and 3600 fall in the CPML and are clamped (three warnings)
    assert gxg.tolist() == [10, 310, 10, 310, 10]
    assert (g.lgest, g.rgest, g.nclamped) == (10, 310, 3)


def test_window_left_of_the_model_and_extension():
    # a shot at 300 m with 1500 m offset: the window starts 1200 m left of the
    # model (mlx < 0); indices are relative to the window, the source at its centre
    gxg, _, g = idx(10, 321, 180, 1500.0, 0.0, 300.0, 0.0, [0.0, 300.0, 1800.0], [0.0] * 3)
    assert g.xsg == 160 and g.mlx == -1200.0
    assert gxg.tolist() == [130, 160, 310]
    # disx widens the window on both sides: the source moves right by disx/dx
    gxg, _, g = idx(10, 331, 180, 1500.0, 50.0, 2000.0, 0.0, [2000.0, 500.0], [0.0] * 2)
    assert g.xsg == 165 and gxg.tolist() == [165, 15]
    assert g.mlx == 450.0


def test_fractional_offset_and_spacing():
    # (1234.5 + 0)/10 + 0.5 = 123.95 -> 123; dz = 7.5: 16/7.5 + 0.5 = 2.63 -> 2
    gxg, gzg, g = idx(8, 263, 60, 1234.5, 0.0, 5000.0, 16.0, [5000.0, 5012.4, 4987.6], [16.0, 3.7, 3.8],
                      dx=10.0, dz=7.5)
    assert (g.xsg, g.zsg) == (131, 10)
    # 12.4/10 + 0.5 = 1.74 -> 1; 3.7/7.5 + 0.5 = 0.993 -> 0; 3.8/7.5 + 0.5 = 1.007 -> 1
    assert gxg.tolist() == [131, 132, 130]
    assert gzg.tolist() == [10, 8, 9]


def test_duplicate_receiver_cells_keep_both_indices():
    # two receivers on one cell keep their own indices (the RTM injection map
    # then lets the later one win, as rindex does)
    gxg, gzg, g = idx(10, 321, 180, 1500.0, 0.0, 2000.0, 0.0, [2100.0, 2101.0, 2099.0], [0.0] * 3)
    assert gxg.tolist() == [170, 170, 170]
    assert (g.lgest, g.rgest) == (170, 170)


def test_source_outside_the_window_is_an_error():
    # xsg = 1500/10 + 0.5 -> 150 + 10 = 160 >= nxx = 100
    with pytest.raises(ValueError):
        idx(10, 100, 180, 1500.0, 0.0, 2000.0, 0.0, [2000.0], [0.0])
    # zsg = 2000/10 + 0.5 -> 200 + 10 >= nzz = 180
    with pytest.raises(ValueError):
        idx(10, 321, 180, 1500.0, 0.0, 2000.0, 2000.0, [2000.0], [0.0])
