"""Face-list mesh generator and the oracle's face/cell passes: invariants a
finite-volume step must keep (no reference test covers this path, so these
are the pins: free-stream preservation, discrete conservation, axis
permutation symmetry of a shock tube, lagged-dt bookkeeping)."""
import numpy as np
import pytest

import oracle
from set_amd import mesh


@pytest.mark.parametrize("dims", [(7, 5, 1), (5, 4, 3)])
def test_face_mesh_topology(dims):
    g = mesh.Grid(*dims) if dims[2] == 1 else mesh.grid_3d(*dims)
    fm = mesh.face_mesh(g)
    nx, ny, nz = dims
    nslot = 4 if nz == 1 else 6
    # every face is listed by its owner; interior faces by both cells
    seen = np.zeros(fm.n_gran, int)
    for c in range(fm.n_cell):
        for s in range(6):
            f = fm.cell_gran[c, s]
            if s >= nslot:
                assert f == 0
                continue
            seen[f - 1] += 1
            s1, s2 = fm.gran_neighbour[f - 1]
            assert c + 1 in (s1, s2)
            # outward normal of the owner points away from the cell centre
            nrm = fm.gran_normal[f - 1]
            d = fm.gran_center[f - 1] - fm.cell_center[c]
            if s1 == c + 1:
                assert d @ nrm > 0
    interior = fm.gran_neighbour[:, 1] > 0
    assert np.all(seen[interior] == 2) and np.all(seen[~interior] == 1)
    assert np.all(fm.gran_type[interior] == mesh.TYPE_INTERIOR)
    assert np.all(fm.gran_type[~interior] == mesh.TYPE_OUTER)
    vol = (g.x1 - g.x0) * (g.y1 - g.y0) * ((g.z1 - g.z0) if nz > 1 else 1.0)
    assert fm.cell_volume.sum() == pytest.approx(vol)
    np.testing.assert_allclose(np.linalg.norm(fm.gran_normal, axis=1), 1.0)
    # TVD neighbours lie beyond s1 / s2 along the normal
    for f in np.nonzero(fm.gran_neighbour_tvd[:, 0] > 0)[0]:
        s1 = fm.gran_neighbour[f, 0]
        t1 = fm.gran_neighbour_tvd[f, 0]
        d = fm.cell_center[s1 - 1] - fm.cell_center[t1 - 1]
        assert d @ fm.gran_normal[f] > 0


def test_free_stream_preserved():
    g = mesh.Grid(24, 12)
    fm = mesh.face_mesh(g)
    W = np.zeros((fm.n_cell, 16))
    W[:, :8] = [0.7, 0.3, -0.2, 0.1, 0.9, 0.5, -0.4, 0.3]
    for cfg in (oracle.cartesian_config(), oracle.cartesian_config(riemann=2, time_scheme=1),
                oracle.cartesian_config(riemann=1)):
        cfg.r_hold_cell = -1.0
        s = oracle.Solver(fm, W, cfg)
        s.dt_next = s.dt_scan()
        assert s.run(5) == 0
        # the flux sum is exactly zero; only the primitive -> conserved ->
        # primitive round trip of the reference's update rounds
        np.testing.assert_allclose(s.W, W, rtol=4e-16, atol=0)


def test_discrete_conservation():
    """Sum of conserved change over the cells = -dt * sum of boundary-face
    fluxes (no sources, no held cells, forward Euler)."""
    g = mesh.Grid(20, 10)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    cfg = oracle.cartesian_config(r_hold=-1.0)
    s = oracle.Solver(fm, W, cfg)
    U0 = mesh.prim_to_cons(s.W[:, :8]) * fm.cell_volume[:, None]
    dt = s.dt_scan()
    s.dt_next = dt
    st, _ = s.face_pass(0)
    assert st == 0
    bnd = s.fm.gran_neighbour[:, 1] < 0
    out = s.potok[bnd, :8].sum(0)  # outward flux x area through the boundary
    assert s.cell_pass(0, dt) == 0
    U1 = mesh.prim_to_cons(s.W[:, :8]) * fm.cell_volume[:, None]
    dU = (U1 - U0).sum(0)
    expect = -dt * out
    expect[5:8] /= mesh.SPI4  # prim_to_cons stores b = B / sqrt(4 pi)
    np.testing.assert_allclose(dU, expect, rtol=1e-9, atol=1e-13 * np.abs(U0).sum(0).max())


def brio_wu(n, axis, nz=1):
    """Brio-Wu shock tube along `axis` on an n x 4 (x 4) grid."""
    shape = [4, 4, 4 if nz > 1 else 1]
    shape[axis] = n
    g = mesh.Grid(shape[0], shape[1], shape[2], 0, 1, 0, 1, 0, 1)
    fm = mesh.face_mesh(g)
    c = fm.cell_center[:, axis]
    W = np.zeros((fm.n_cell, 16))
    left = c < 0.5
    W[:, 0] = np.where(left, 1.0, 0.125)
    W[:, 4] = np.where(left, 1.0, 0.1)
    t = [a for a in range(3) if a != axis]
    W[:, 5 + axis] = 0.75 * mesh.SPI4
    W[:, 5 + t[0]] = np.where(left, 1.0, -1.0) * mesh.SPI4
    return g, fm, W, c


@pytest.mark.parametrize("riemann", [2, 3])
def test_shock_tube_axis_permutation(riemann):
    res = []
    for axis in (0, 1, 2):
        g, fm, W, c = brio_wu(64, axis, nz=3 if axis == 2 else 1)
        cfg = oracle.cartesian_config(riemann=riemann, r_hold=-1.0)
        s = oracle.Solver(fm, W, cfg)
        s.dt_next = s.dt_scan()
        assert s.run(30) == 0
        order = np.argsort(c, kind="stable")
        prof = s.W[order][:, [0, 4]].reshape(64, -1, 2)[:, 0, :]
        res.append(prof)
    np.testing.assert_allclose(res[0], res[1], rtol=1e-12)
    np.testing.assert_allclose(res[0], res[2], rtol=1e-12)
    assert np.all(np.isfinite(res[0])) and np.all(res[0] > 0)
    # the left state is untouched beyond the fastest wave's reach
    assert res[0][0, 0] == 1.0 and res[0][-1, 0] == 0.125


def test_lagged_dt_trace():
    """dt of step 1 is the initial dt; dt of step 2 is the CFL minimum of
    the state step 1 started from (cuf_kernel.cuf:789-790)."""
    g = mesh.Grid(30, 15)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g)
    cfg = oracle.cartesian_config()
    s = oracle.Solver(fm, W, cfg)
    dt0 = s.dt_scan()
    s.dt_next = 1.0e-5
    s.run(1)
    assert s.t == 1.0e-5
    assert s.dt_next == dt0
    s.run(1)
    assert s.t == 1.0e-5 + dt0
