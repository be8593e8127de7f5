"""FCMHD binary format (Storage_module.f90:43-209, Setka.cpp:1240-1565):
round trips, Set_Storage's rejections, and the byte layout."""
import numpy as np
import pytest

from set_amd import fcmhd, mesh


def small():
    g = mesh.grid_2d(6, 4)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g)
    W[:, 8:] = np.arange(fm.n_cell * 8).reshape(fm.n_cell, 8) * 0.01 + 0.1
    return fm, W


def test_input_round_trip(tmp_path):
    fm, W = small()
    p = tmp_path / "FCMHD_3.bin"
    pot = np.random.default_rng(0).normal(size=(fm.n_gran, 9))
    fcmhd.write_input(p, fm, W, time=1.5, gran_potok=pot)
    t, fm2, W2, pot2 = fcmhd.read_input(p)
    assert t == 1.5
    np.testing.assert_array_equal(W2, W)
    np.testing.assert_array_equal(pot2, pot)
    for k in ("cell_center", "cell_volume", "cell_gran", "gran_normal", "gran_square",
              "gran_center", "gran_neighbour", "gran_neighbour_tvd", "gran_type"):
        np.testing.assert_array_equal(getattr(fm2, k), getattr(fm, k))
    # byte layout: header 16 B, then arrays, three sentinels
    nc, nf = fm.n_cell, fm.n_gran
    size = 8 + 8 + 8 * (16 + 3 + 1) * nc + 4 * 6 * nc + 8 * (3 + 1 + 3) * nf + 4 * 5 * nf + 8 * 9 * nf + 24
    assert p.stat().st_size == size


def test_rejections(tmp_path):
    fm, W = small()
    p = tmp_path / "bad.bin"
    bad = W.copy()
    bad[3, 0] = 1e-12
    fcmhd.write_input(p, fm, bad)
    with pytest.raises(fcmhd.FCMHDError, match="rho"):
        fcmhd.read_input(p)
    bad = W.copy()
    bad[2, 11] = -1.0
    fcmhd.write_input(p, fm, bad)
    with pytest.raises(fcmhd.FCMHDError, match="12"):
        fcmhd.read_input(p)
    fcmhd.write_input(p, fm, W)
    raw = bytearray(p.read_bytes())
    raw[-8:] = np.array([124.0]).tobytes()
    p.write_bytes(bytes(raw))
    with pytest.raises(fcmhd.FCMHDError, match="123"):
        fcmhd.read_input(p)
    p.write_bytes(bytes(raw[:-30]))
    with pytest.raises(fcmhd.FCMHDError, match="ends"):
        fcmhd.read_input(p)


def test_save_fill_round_trip(tmp_path):
    fm, W = small()
    p = tmp_path / "FCMHD_3.5_out.bin"
    fcmhd.save_storage(p, 2.25, W)
    assert p.stat().st_size == 8 + 8 * 16 * fm.n_cell + 8
    t, W2 = fcmhd.fill_data(p, fm.n_cell)
    assert t == 2.25
    np.testing.assert_array_equal(W2, W)
