"""Pins of the oracle's face pass -- minmod TVD reconstruction, its
positivity fallback, the boundary ghosts of Gran_type 1/3/4/5 and the wall
(type 2 with no second cell) -- and of Source_MK.

Each test checks the oracle against a statement that does not share its
text: the textbook minmod-MUSCL reconstruction written in coordinates along
the grid line, the exact reproduction of linear profiles, the ideal-MHD
physical flux, mirror images of the domain, and invariances and special
values of the charge-exchange source.  chlld itself is pinned bit for bit
to the reference's compiled routine (test_oracle.py), so where a face's
Riemann problem is needed these tests call the oracle's chlld and check
what the face pass feeds it.  Reference lines: cuf_solvers.cuf:318-412
(ghosts, TVD), :414-420 (face density STOP), cuf_kernel.cuf:215-288
(Source_MK)."""
import numpy as np
import pytest

import oracle
from set_amd import mesh
from test_oracle import phys_flux

F4 = oracle.F4
ORC_OK, ORC_ERR_FACE_RHO = 0, -3


def tensor_mesh(xe, ye):
    """SET's face-list arrays for the 2-D tensor-product grid with cell edges
    xe (x) and ye (y): the uniform grid's connectivity with the geometry of
    the given edges (cell centres at edge midpoints)."""
    nx, ny = len(xe) - 1, len(ye) - 1
    fm = mesh.face_mesh(mesh.Grid(nx, ny, 1, xe[0], xe[-1], ye[0], ye[-1]))
    xc, yc = 0.5 * (xe[1:] + xe[:-1]), 0.5 * (ye[1:] + ye[:-1])
    dx, dy = np.diff(xe), np.diff(ye)
    c = np.arange(fm.n_cell)
    i, j = c % nx, c // nx
    fm.cell_center[:, 0], fm.cell_center[:, 1] = xc[i], yc[j]
    fm.cell_volume[:] = dx[i] * dy[j]
    nfx = (nx + 1) * ny
    f = np.arange(nfx)
    fi, fj = f % (nx + 1), f // (nx + 1)
    fm.gran_center[:nfx, 0], fm.gran_center[:nfx, 1], fm.gran_square[:nfx] = xe[fi], yc[fj], dy[fj]
    f = np.arange(fm.n_gran - nfx)
    fi, fj = f % nx, f // nx
    fm.gran_center[nfx:, 0], fm.gran_center[nfx:, 1], fm.gran_square[nfx:] = xc[fi], ye[fj], dx[fi]
    return fm


def stretched(rng, nx, ny):
    wx = rng.uniform(0.4, 1.6, nx)
    wy = rng.uniform(0.4, 1.6, ny)
    return tensor_mesh(np.r_[0.0, np.cumsum(wx)] / wx.sum(), np.r_[0.0, np.cumsum(wy)] / wy.sum())


def minmod(a, b):
    return np.where(a * b <= 0.0, 0.0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)))


def muscl_face_states(fm, W, f):
    """Textbook minmod-MUSCL states of interior face f (0-based) in
    coordinates X along the face's grid line: the left cell's value plus the
    limited slope times (x_face - X_left), the right cell's likewise.  rho,
    u, v, w, p are reconstructed, B stays first order; a negative rho or p
    falls back to the cell's value.  Faces missing a second neighbour on
    either side are first order."""
    s1, s2 = fm.gran_neighbour[f] - 1
    t1, t2 = fm.gran_neighbour_tvd[f] - 1
    qL, qR = W[s1, :8].copy(), W[s2, :8].copy()
    if t1 < 0 or t2 < 0:
        return qL, qR
    ax = int(np.argmax(np.abs(fm.gran_normal[f])))
    X = fm.cell_center[:, ax]
    xf = fm.gran_center[f, ax]
    a, b, c, d = W[t1, :5], W[s1, :5], W[s2, :5], W[t2, :5]
    xa, xb, xc, xd = X[t1], X[s1], X[s2], X[t2]
    sc = (c - b) / (xc - xb)
    left = b + minmod((b - a) / (xb - xa), sc) * (xf - xb)
    right = c + minmod(sc, (d - c) / (xd - xc)) * (xf - xc)
    for v in (0, 4):
        if left[v] < 0.0:
            left[v] = b[v]
        if right[v] < 0.0:
            right[v] = c[v]
    qL[:5], qR[:5] = left, right
    return qL, qR


def face_flux(cfg, fm, W):
    cfg.r_hold_cell = -1.0
    s = oracle.Solver(fm, W, cfg)
    st, _ = s.face_pass(0)
    assert st == 0
    return s.potok


def random_state(rng, n, b=0.6):
    W = np.zeros((n, 16))
    W[:, 0] = rng.uniform(0.5, 2.0, n)
    W[:, 1:4] = rng.normal(size=(n, 3)) * 0.5
    W[:, 4] = rng.uniform(0.5, 2.0, n)
    W[:, 5:8] = rng.normal(size=(n, 3)) * b
    return W


@pytest.mark.parametrize("riemann", [1, 2, 3])
def test_tvd_states_are_minmod_muscl(riemann):
    """Every interior face's flux is chlld of the textbook MUSCL states, on a
    grid stretched in both axes (a swapped distance, TVD neighbour or side
    changes the states there)."""
    rng = np.random.default_rng(31 + riemann)
    fm = stretched(rng, 9, 8)
    W = random_state(rng, fm.n_cell)
    pot = face_flux(oracle.cartesian_config(riemann=riemann, recon=1), fm, W)
    n_tvd = 0
    for f in range(fm.n_gran):
        if fm.gran_neighbour[f, 1] < 0:
            continue
        qL, qR = muscl_face_states(fm, W, f)
        n_tvd += fm.gran_neighbour_tvd[f].min() > 0
        ref = oracle.chlld(riemann, fm.gran_normal[f], qL, qR)[0] * fm.gran_square[f]
        np.testing.assert_allclose(pot[f, :8], ref, rtol=0, atol=1e-13 * np.abs(ref).max(), err_msg=f"face {f}")
    assert n_tvd > 50


def test_tvd_positivity_fallback():
    """The limited value lies between the two cells' values, so a face
    pressure or density below zero needs a negative neighbour; the face's
    Riemann problem is then non-finite either way, and the fallback
    (cuf_solvers.cuf:398-412) shows only in the face states.  A neighbour
    of negative pressure drives the limited pressure of its neighbour's side
    below zero: that side takes its cell's own pressure, the other
    variables keep their MUSCL values.  Likewise for density, where the
    negative cell then stops the face pass (:414-420)."""
    rng = np.random.default_rng(40)
    wy = rng.uniform(0.4, 1.6, 6)
    fm = tensor_mesh(np.linspace(0.0, 1.0, 9), np.r_[0.0, np.cumsum(wy)] / wy.sum())
    nx = 8
    col = np.arange(fm.n_cell) % nx
    for v, bad in ((4, ORC_OK), (0, ORC_ERR_FACE_RHO)):
        W = random_state(rng, fm.n_cell)
        W[col == 3, v] = 4.0
        W[col == 4, v] = 1.0
        W[col == 5, v] = -2.0
        s = oracle.Solver(fm, W, oracle.cartesian_config(riemann=1, recon=1, r_hold=-1.0))
        hit = 0
        for f in range(fm.n_gran):
            s1, s2 = fm.gran_neighbour[f] - 1
            if s2 < 0 or fm.gran_normal[f, 0] != 1.0 or col[s1] != 4:
                continue
            st, q1, q2 = s.face_states(f)
            assert st == bad
            qL, qR = muscl_face_states(fm, W, f)
            assert qL[v] == W[s1, v] == q1[v] == 1.0  # the fallback fired
            np.testing.assert_allclose(q1, qL, rtol=1e-14, atol=1e-15)
            hit += 1
        assert hit == 6


def test_negative_face_density_stops():
    """A cell of negative density makes the face pass report the reference's
    face-density STOP (cuf_solvers.cuf:414-420)."""
    g = mesh.Grid(8, 4)
    fm = mesh.face_mesh(g)
    W = random_state(np.random.default_rng(41), fm.n_cell)
    W[10, 0] = -0.5
    s = oracle.Solver(fm, W, oracle.cartesian_config(recon=1, r_hold=-1.0))
    st, _ = s.face_pass(0)
    assert st == ORC_ERR_FACE_RHO


@pytest.mark.parametrize("riemann", [0, 2, 3])
def test_tvd_reproduces_linear_profiles(riemann):
    """rho, u, v, w, p linear in x and y, B uniform, on a stretched grid: the
    limited reconstruction is exact, both sides of a face with two TVD
    neighbours carry the profile's value at the face centre, and chlld of
    equal states is the physical flux."""
    rng = np.random.default_rng(42)
    fm = stretched(rng, 10, 9)
    a = np.array([1.0, 0.3, -0.2, 0.1, 1.2])
    gx = np.array([0.4, 0.5, -0.3, 0.2, 0.6])
    gy = np.array([-0.3, 0.2, 0.4, -0.1, 0.5])
    B = np.array([0.5, -0.3, 0.2])

    def prof(xy):
        q = np.zeros(8)
        q[:5] = a + gx * xy[0] + gy * xy[1]
        q[5:] = B
        return q

    W = np.zeros((fm.n_cell, 16))
    for c in range(fm.n_cell):
        W[c, :8] = prof(fm.cell_center[c])
    pot = face_flux(oracle.cartesian_config(riemann=riemann, recon=1), fm, W)
    n = 0
    for f in range(fm.n_gran):
        if fm.gran_neighbour[f, 1] < 0 or fm.gran_neighbour_tvd[f].min() < 1:
            continue
        ref = phys_flux(prof(fm.gran_center[f]), fm.gran_normal[f]) * fm.gran_square[f]
        np.testing.assert_allclose(pot[f, :8], ref, rtol=0, atol=2e-12 * np.abs(ref).max(), err_msg=f"face {f}")
        n += 1
    assert n > 80


def test_ghost_states():
    """The ghost each boundary type hands to chlld: types 4, 1, 3 a copy of
    the cell (4 with its x-velocity raised to -0.1 under the heliosphere
    switches), type 5 the cell with v negated, the wall the cell with its
    velocity reflected in the face, v - 2 (v.n) n; no TVD on any of them."""
    g = mesh.Grid(6, 5, 1, -1.0, 0.0, 0.0, 1.0)
    rng = np.random.default_rng(48)
    for ftype in (4, 1, 3, 5, 2):
        fm = mesh.face_mesh(g)
        bnd = np.nonzero(fm.gran_neighbour[:, 1] < 0)[0]
        n = rng.normal(size=(len(bnd), 3))
        fm.gran_normal[bnd] = n / np.linalg.norm(n, axis=1)[:, None]
        fm.gran_type[bnd] = ftype
        W = random_state(rng, fm.n_cell)
        W[::2, 1] = -0.05
        for hacks in (0, oracle.F_HELIO_HACKS):
            cfg = oracle.cartesian_config(riemann=2, recon=1, r_hold=-1.0)
            cfg.flags = hacks
            s = oracle.Solver(fm, W, cfg)
            for f in bnd:
                st, q1, q2 = s.face_states(f)
                q = W[fm.gran_neighbour[f, 0] - 1, :8]
                assert st == 0
                np.testing.assert_array_equal(q1, q)
                want = q.copy()
                if ftype == 4 and hacks and q[1] > F4(-0.1):
                    want[1] = F4(-0.1)
                elif ftype == 5:
                    want[2] = -q[2]
                elif ftype == 2:
                    nn = fm.gran_normal[f]
                    want[1:4] = q[1:4] - 2.0 * (q[1:4] @ nn) * nn
                np.testing.assert_allclose(q2, want, rtol=0, atol=1e-15 * np.abs(q).max())


@pytest.mark.parametrize("ghost_type", [4, 1, 3])
def test_copy_ghosts_give_physical_flux(ghost_type):
    """Outer boundary faces (type 4 without the heliosphere clamp) and faces
    of types 1 and 3 without a second cell take a copy of their cell: the
    flux is the cell's physical flux through the outward normal, first
    order also under TVD."""
    rng = np.random.default_rng(43)
    fm = stretched(rng, 7, 6)
    bnd = fm.gran_neighbour[:, 1] < 0
    fm.gran_type[bnd] = ghost_type
    W = random_state(rng, fm.n_cell)
    for riemann in (1, 2, 3):
        pot = face_flux(oracle.cartesian_config(riemann=riemann, recon=1), fm, W)
        for f in np.nonzero(bnd)[0]:
            q = W[fm.gran_neighbour[f, 0] - 1, :8]
            ref = phys_flux(q, fm.gran_normal[f]) * fm.gran_square[f]
            np.testing.assert_allclose(pot[f, :8], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_outer_clamp_u():
    """Type 4 with the heliosphere switches (cuf_solvers.cuf:318-324): the
    ghost's x-velocity is raised to no more than -0.1 whatever the normal;
    a cell already at u <= -0.1 keeps a copy ghost."""
    g = mesh.Grid(6, 4, 1, -1.0, 0.0, 0.0, 1.0)  # x < 0.22: no LAX wedge face
    fm = mesh.face_mesh(g)
    bnd = np.nonzero(fm.gran_neighbour[:, 1] < 0)[0]
    rng = np.random.default_rng(44)
    W = random_state(rng, fm.n_cell, b=0.0)
    W[:, 1] = np.where(np.arange(fm.n_cell) % 2 == 0, -0.05, -0.3)
    cfg = oracle.cartesian_config(riemann=2, recon=1)
    cfg.flags = oracle.F_HELIO_HACKS | oracle.F_ZERO_BW
    cfg.r_skip_face = -1.0
    pot = face_flux(cfg, fm, W)
    clamped = 0
    for f in bnd:
        q = W[fm.gran_neighbour[f, 0] - 1, :8]
        g2 = q.copy()
        if q[1] > F4(-0.1):
            g2[1] = F4(-0.1)
            clamped += 1
            assert q[1] == -0.05
        ref = oracle.chlld(2, fm.gran_normal[f], q, g2)[0] * fm.gran_square[f]
        np.testing.assert_array_equal(pot[f, :8], ref)
        if q[1] == -0.3:
            np.testing.assert_allclose(pot[f, :8], phys_flux(q, fm.gran_normal[f]) * fm.gran_square[f],
                                       rtol=0, atol=1e-12 * np.abs(ref).max())
    assert 0 < clamped < len(bnd)


def hydro_state(rng, n):
    W = random_state(rng, n, b=0.0)
    W[:, 1:3] *= 0.6
    W[:, 3] = 0.0
    return W


@pytest.mark.parametrize("riemann,scheme", [(2, 0), (3, 1), (1, 0), (0, 1)])
def test_axis_ghost_is_mirror_image(riemann, scheme):
    """Type 5 (cuf_solvers.cuf:325-329): the half plane y > 0 with axis faces
    at y = 0 evolves as the upper half of the full plane holding the mirror
    image (v -> -v) below the axis.  Only the axis faces' rounding differs
    (chlld is solved in mirrored frames), so to 1e-13 after 10 steps."""
    nx, ny = 16, 8  # spacing 1/8: cell centres are exact in both meshes
    rng = np.random.default_rng(45)
    half = mesh.face_mesh(mesh.Grid(nx, ny, 1, -1.0, 1.0, 0.0, 1.0))
    axis = (half.gran_normal[:, 1] == -1.0) & (half.gran_center[:, 1] == 0.0)
    half.gran_type[axis] = mesh.TYPE_AXIS
    full = mesh.face_mesh(mesh.Grid(nx, 2 * ny, 1, -1.0, 1.0, -1.0, 1.0))
    Wh = hydro_state(rng, half.n_cell)
    Wf = np.zeros((full.n_cell, 16))
    j = np.arange(full.n_cell) // nx
    i = np.arange(full.n_cell) % nx
    src = np.where(j >= ny, j - ny, ny - 1 - j) * nx + i
    Wf[:] = Wh[src]
    Wf[j < ny, 2] *= -1.0
    cfg = oracle.cartesian_config(riemann=riemann, recon=0, time_scheme=scheme, r_hold=-1.0)
    sh, sf = oracle.Solver(half, Wh, cfg), oracle.Solver(full, Wf, cfg)
    dt0 = sh.dt_scan()
    assert sf.dt_scan() == dt0
    sh.dt_next = sf.dt_next = dt0
    assert sh.run(10) == 0 and sf.run(10) == 0
    up = sf.W[j >= ny, :8]
    scale = np.abs(sh.W[:, :8]).max(0)
    assert np.all(np.abs(up - sh.W[:, :8]).max(0) <= 1e-13 * np.maximum(scale, 1.0))
    assert sh.t == pytest.approx(sf.t, rel=1e-13)
    # the lower half stays the mirror image
    lo = sf.W[j < ny, :8]
    lo_src = sf.W[j >= ny][(ny - 1 - j[j < ny]) * nx + i[j < ny]][:, :8]
    mirror = lo_src.copy()
    mirror[:, 2] *= -1.0
    assert np.all(np.abs(lo - mirror).max(0) <= 1e-13 * np.maximum(scale, 1.0))


@pytest.mark.parametrize("riemann,scheme", [(2, 1), (3, 0), (1, 1), (0, 0)])
def test_wall_is_mirror_image(riemann, scheme):
    """Type 2 with no second cell is a reflecting wall (cuf_solvers.cuf:
    330-345).  The box [0,1]^2 with a wall at x = 1 evolves as the left half
    of [0,2] x [0,1] holding the mirror image (u -> -u) beyond x = 1: the
    wall's ghost state is exactly the mirror cell's state, so one Euler
    step agrees bit for bit (later the mirror half drifts from an exact
    image by rounding: 1e-13 after 10 steps).  No mass crosses the wall."""
    nx, ny = 8, 8
    rng = np.random.default_rng(46)
    box = mesh.face_mesh(mesh.Grid(nx, ny, 1, 0.0, 1.0, 0.0, 1.0))
    wall = (box.gran_normal[:, 0] == 1.0) & (box.gran_neighbour[:, 1] < 0)
    box.gran_type[wall] = mesh.TYPE_INTERIOR  # type 2, s2 = -1
    full = mesh.face_mesh(mesh.Grid(2 * nx, ny, 1, 0.0, 2.0, 0.0, 1.0))
    Wb = hydro_state(rng, box.n_cell)
    i = np.arange(full.n_cell) % (2 * nx)
    j = np.arange(full.n_cell) // (2 * nx)
    src = np.where(i < nx, i, 2 * nx - 1 - i) + nx * j
    Wf = Wb[src].copy()
    Wf[i >= nx, 1] *= -1.0
    cfg = oracle.cartesian_config(riemann=riemann, recon=0, time_scheme=scheme, r_hold=-1.0)
    sb, sf = oracle.Solver(box, Wb, cfg), oracle.Solver(full, Wf, cfg)
    dt0 = sb.dt_scan()
    assert sf.dt_scan() == dt0
    sb.dt_next = sf.dt_next = dt0
    if scheme == 0:
        assert sb.run(1) == 0 and sf.run(1) == 0
        np.testing.assert_array_equal(sf.W[i < nx, :8], sb.W[:, :8])
        assert sb.dt_next == sf.dt_next
    assert sb.run(10) == 0 and sf.run(10) == 0
    scale = np.maximum(np.abs(sb.W[:, :8]).max(0), 1.0)
    assert np.all(np.abs(sf.W[i < nx, :8] - sb.W[:, :8]).max(0) <= 1e-13 * scale)
    assert sb.t == pytest.approx(sf.t, rel=1e-13)
    st, _ = sb.face_pass(0)
    assert st == 0
    rho_c = np.abs(sb.W[:, 0] * (np.abs(sb.W[:, 1]) + np.sqrt(sb.W[:, 4] / sb.W[:, 0]))).max()
    assert np.abs(sb.potok[wall, 0]).max() <= 1e-13 * rho_c * box.gran_square[wall].max()


def test_wall_oblique_normal():
    """The wall reflects the normal velocity for any normal: the ghost
    velocity is v - 2 (v.n) n, so the flux equals chlld of the cell and that
    reflected state (checked against the reflection written in numpy)."""
    g = mesh.Grid(6, 5)
    fm = mesh.face_mesh(g)
    rng = np.random.default_rng(47)
    bnd = np.nonzero(fm.gran_neighbour[:, 1] < 0)[0]
    n = rng.normal(size=(len(bnd), 3))
    fm.gran_normal[bnd] = n / np.linalg.norm(n, axis=1)[:, None]
    fm.gran_type[bnd] = mesh.TYPE_INTERIOR
    W = random_state(rng, fm.n_cell)
    pot = face_flux(oracle.cartesian_config(riemann=3, recon=1), fm, W)
    for f in bnd:
        q = W[fm.gran_neighbour[f, 0] - 1, :8]
        nn = fm.gran_normal[f]
        g2 = q.copy()
        g2[1:4] = q[1:4] - 2.0 * (q[1:4] @ nn) * nn
        ref = oracle.chlld(3, nn, q, g2)[0] * fm.gran_square[f]
        np.testing.assert_allclose(pot[f, :8], ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("dim", [2, 3])
def test_free_stream_preserved_tvd(scheme, dim):
    g = mesh.Grid(16, 10) if dim == 2 else mesh.grid_3d(8, 7, 6)
    fm = mesh.face_mesh(g)
    W = np.zeros((fm.n_cell, 16))
    W[:, :8] = [0.7, 0.3, -0.2, 0.1, 0.9, 0.5, -0.4, 0.3]
    for riemann in (2, 3):
        cfg = oracle.cartesian_config(riemann=riemann, recon=1, time_scheme=scheme, r_hold=-1.0)
        s = oracle.Solver(fm, W, cfg)
        s.dt_next = s.dt_scan()
        assert s.run(4) == 0
        np.testing.assert_allclose(s.W, W, rtol=4e-16, atol=0)


def test_conservation_tvd_heun():
    """Minmod TVD + Heun, stage by stage: stage 1 changes the conserved
    totals by -dt times the boundary flux of U^n; stage 2 makes them
    (total(U^n) + total(U^1))/2 - dt/2 times the boundary flux of U^1."""
    g = mesh.Grid(20, 12)
    fm = mesh.face_mesh(g)
    # a warm wind: no cell reaches the pressure floor, which adds energy
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3, speed=2.0, p_over_rho=0.5))
    cfg = oracle.cartesian_config(recon=1, time_scheme=1, r_hold=-1.0)
    s = oracle.Solver(fm, W, cfg)
    vol = fm.cell_volume[:, None]
    bnd = fm.gran_neighbour[:, 1] < 0

    def total(Wx):
        return (mesh.prim_to_cons(Wx[:, :8]) * vol).sum(0)

    def bflux():
        out = s.potok[bnd, :8].sum(0)
        out[5:8] /= mesh.SPI4
        return out

    T0 = total(s.W)
    dt = s.dt_scan()
    st, _ = s.face_pass(1)
    assert st == 0
    B0 = bflux()
    assert s.cell_pass(1, dt) == 0
    T1 = total(s.W2)
    scale = np.abs(mesh.prim_to_cons(W[:, :8]) * vol).sum(0).max()
    np.testing.assert_allclose(T1 - T0, -dt * B0, rtol=1e-9, atol=1e-13 * scale)
    st, _ = s.face_pass(2)
    assert st == 0
    B1 = bflux()
    assert s.cell_pass(2, dt) == 0
    T2 = total(s.W)
    np.testing.assert_allclose(T2 - 0.5 * (T0 + T1), -0.5 * dt * B1, rtol=1e-9, atol=1e-13 * scale)


def test_six_face_slots_are_summed():
    """The reference sums Cell_gran slots 1..4 (cuf_solvers.cuf:691); SET's
    meshes hold 0 in slots 5 and 6.  The oracle sums all six so that 3-D
    face-list meshes can place their z-faces there: a 2-D mesh is unchanged
    by it, and on a 3-D mesh dropping slots 5 and 6 loses the z-fluxes."""
    g = mesh.grid_3d(6, 5, 4)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.5, speed=2.0))
    cfg = oracle.cartesian_config(r_hold=-1.0)
    s6 = oracle.Solver(fm, W, cfg)
    dt = s6.dt_scan()
    s6.dt_next = dt
    assert s6.run(1) == 0
    fm4 = mesh.face_mesh(g)
    fm4.cell_gran[:, 4:6] = 0
    s4 = oracle.Solver(fm4, W, cfg)
    s4.dt_next = dt
    assert s4.run(1) == 0
    assert np.abs(s4.W[:, 3] - s6.W[:, 3]).max() > 1e-6
    # 2-D: slots 5, 6 are 0 and the sum is the reference's four-slot sum
    fm2 = mesh.face_mesh(mesh.Grid(8, 6))
    assert np.all(fm2.cell_gran[:, 4:6] == 0)


# ---------------------------------------------------------------------------
# Source_MK (cuf_kernel.cuf:215-288)

def mk_state(rng):
    return np.r_[rng.uniform(0.1, 2), rng.normal(size=3), rng.uniform(0.1, 2), np.zeros(3),
                 rng.uniform(0.01, 1), rng.normal(size=2), rng.uniform(0.1, 2),
                 rng.uniform(0.01, 1), rng.normal(size=2), rng.uniform(0.1, 2)]


def test_source_mk_galilean():
    """Shifting the plasma and both neutral populations by a common velocity
    c leaves the momentum exchange unchanged and adds c . (momentum
    exchange) to the energy exchange (a dropped or mis-signed kinetic term
    breaks it)."""
    rng = np.random.default_rng(50)
    for _ in range(100):
        cp = mk_state(rng)
        S = oracle.source_mk(cp)
        cx, cy = rng.normal(size=2) * 2
        sh = cp.copy()
        sh[[1, 9, 13]] += cx
        sh[[2, 10, 14]] += cy
        T = oracle.source_mk(sh)
        scale = np.abs(S).max() * (1 + abs(cx) + abs(cy))
        assert abs(T[0] - S[0]) <= 1e-12 * scale and abs(T[1] - S[1]) <= 1e-12 * scale
        assert abs(T[2] - (S[2] + cx * S[0] + cy * S[1])) <= 1e-11 * scale


def test_source_mk_population_swap():
    """The two neutral populations enter alike: swapping H3 and H4 gives the
    same source bit for bit (addition commutes)."""
    rng = np.random.default_rng(51)
    for _ in range(100):
        cp = mk_state(rng)
        sw = cp.copy()
        sw[8:12], sw[12:16] = cp[12:16], cp[8:12]
        np.testing.assert_array_equal(oracle.source_mk(sw), oracle.source_mk(cp))


def test_source_mk_density_floor():
    """A neutral density below 1e-7 (a default-REAL literal) is raised to it
    and its pressure dropped (cuf_kernel.cuf:242-254); at 2e-7 nothing is
    floored."""
    rng = np.random.default_rng(52)
    for _ in range(50):
        cp = mk_state(rng)
        lo = cp.copy()
        lo[8], lo[11] = 1e-9, 5.0
        ref = cp.copy()
        ref[8], ref[11] = F4(1e-7), 0.0
        np.testing.assert_array_equal(oracle.source_mk(lo), oracle.source_mk(ref))
        hi = cp.copy()
        hi[8], hi[11] = 2e-7, 5.0
        hi0 = hi.copy()
        hi0[11] = 0.0
        assert np.any(oracle.source_mk(hi) != oracle.source_mk(hi0))


def test_source_mk_scaling():
    """Doubling the plasma density and pressure (same temperature) doubles
    the source bit for bit: it is linear in the plasma density."""
    rng = np.random.default_rng(53)
    for _ in range(50):
        cp = mk_state(rng)
        d = cp.copy()
        d[[0, 4]] *= 2.0
        np.testing.assert_array_equal(oracle.source_mk(d), 2.0 * oracle.source_mk(cp))


def two_alike(rho, u, v, p, rh, uh, vh, th):
    """Cell_par with both neutral populations equal: the source is then
    exactly twice one population's term (a + a = 2a)."""
    return np.r_[rho, u, v, 0.0, p, 0, 0, 0, rh, uh, vh, th, rh, uh, vh, th]


def test_source_mk_cold_closed_form():
    """Cold plasma and neutrals (p = T_H = 0): the relative speed is |dV|
    and sigma = (1 - 0.1307345665 ln|dV|)^2.  At |dV| = 1, sigma = 1 and the
    momentum exchange is (3/0.4326569) rho rho_H dV per population; at
    |dV| = exp(1/0.1307345665) the cross-section vanishes."""
    kk = float(np.float32(3.0) / np.float32(0.4326569))
    ks = F4(0.1307345665)
    for rho, rh, ang in ((1.3, 0.4, 0.3), (0.2, 1.7, 2.1)):
        du, dv = np.cos(ang), np.sin(ang)
        S = oracle.source_mk(two_alike(rho, 0.2, -0.1, 0.0, rh, 0.2 + du, -0.1 + dv, 0.0)) / 2
        np.testing.assert_allclose(S[:2], kk * rho * rh * np.array([du, dv]), rtol=1e-12)
        big = np.exp(1.0 / ks)
        Z = oracle.source_mk(two_alike(rho, 0.2, -0.1, 0.0, rh, 0.2 + big * du, -0.1 + big * dv, 0.0)) / 2
        assert np.abs(Z[:2]).max() <= 1e-9 * kk * rho * rh * big


def test_source_mk_thermal_closed_form():
    """No drift, thermal speeds with (64/9pi)(p/rho + T_H) = 1 (T_H =
    2 pH/rhoH): U_M = 1, sigma = 1, no momentum exchange, and the energy
    exchange per population is kk rho rho_H (U/U_M)(T_H - p/rho) with
    U/U_M = sqrt((4/pi)/(64/9pi)) = 3/4."""
    kk = float(np.float32(3.0) / np.float32(0.4326569))
    tot = 9.0 * np.pi / 64.0
    for rho, rh, frac in ((1.1, 0.3, 0.2), (0.4, 1.5, 0.7)):
        p = frac * tot * rho
        th = (1.0 - frac) * tot
        S = oracle.source_mk(two_alike(rho, 0.3, -0.4, p, rh, 0.3, -0.4, th)) / 2
        assert np.abs(S[:2]).max() == 0.0
        np.testing.assert_allclose(S[2], kk * rho * rh * 0.75 * (th - p / rho), rtol=1e-12)


# Probes of the heliosphere switches' regions, with the region each lies in
# written out by hand from the reference text (cuf_solvers.cuf:296-298 for the
# LAX wedge, :783-785 for the v clamp).  The thresholds are REAL*4 literals
# compared in REAL*8: 0.22 -> 0.2199999988, 0.1 -> 0.1000000015,
# 0.28 -> 0.2800000012, 0.07 -> 0.0700000003; the comparisons are strict.
# Probes between a literal's single and double value fail a double literal;
# probes on either side of each threshold fail a reversed or non-strict
# comparison, a swapped axis or the wrong normal component.
LAX_PROBES = [  # (face centre x, face centre y, normal x, LAX?)
    (0.5, 0.05, 0.0, True),
    (0.2200001, 0.05, 0.0, True),
    (0.22, 0.05, 0.0, True),             # above 0.2199999988
    (0.2199999995, 0.05, 0.0, True),     # between the single and the double 0.22
    (0.219999998, 0.05, 0.0, False),
    (-0.5, 0.05, 0.0, False),
    (0.5, 0.1000000010, 0.0, True),      # between the double and the single 0.1
    (0.5, 0.1000000020, 0.0, False),
    (0.5, -0.3, 0.0, True),              # no absolute value on y
    (0.05, 0.5, 0.0, False),             # x and y swapped
    (0.5, 0.05, 0.0999999, True),
    (0.5, 0.05, 0.1000000010, True),     # between the double and the single 0.1
    (0.5, 0.05, -0.1000000010, True),    # |n_x|
    (0.5, 0.05, 0.1000000020, False),
    (0.5, 0.05, -0.3, False),
    (float(np.float32(0.22)), 0.05, 0.0, False),   # on each threshold: strict comparisons
    (0.5, float(np.float32(0.1)), 0.0, False),
    (0.5, 0.05, float(np.float32(0.1)), False),
]

CLAMP_PROBES = [  # (cell centre x, cell centre y, v clamped to 0.01?)
    (0.5, 0.03, True),
    (0.2800001, 0.03, True),
    (0.2800000005, 0.03, False),         # above the double 0.28, below the single
    (0.28, 0.03, False),
    (0.5, 0.0700000001, True),           # below the single 0.07, above the double
    (0.5, 0.0700000005, False),
    (0.5, -0.2, True),                   # no absolute value on y
    (0.03, 0.5, False),                  # x and y swapped
    (float(np.float32(0.28)), 0.03, False),  # on each threshold: strict comparisons
    (0.5, float(np.float32(0.07)), False),
]


def helio_cfg(riemann):
    cfg = oracle.cartesian_config(riemann=riemann, recon=0)
    cfg.flags = oracle.F_HELIO_HACKS
    cfg.r_skip_face = -1.0
    cfg.r_hold_cell = -1.0
    return cfg


@pytest.mark.parametrize("riemann", [2, 3])
def test_lax_wedge_region(riemann):
    """With the heliosphere switches a face takes the LAX flux (n_state 0)
    exactly where the reference's wedge test holds, and the configured
    solver elsewhere; each probe face's flux is compared bit for bit with
    chlld of its two cells through its (possibly tilted) normal."""
    rng = np.random.default_rng(61)
    fm = tensor_mesh(np.linspace(0.0, 1.0, 7), np.linspace(0.0, 1.0, 7))
    inner = np.nonzero(fm.gran_neighbour[:, 1] > 0)[0]
    assert len(inner) >= len(LAX_PROBES)
    fm.gran_type[:] = 2
    fm.gran_type[fm.gran_neighbour[:, 1] < 0] = 1
    probes = inner[rng.permutation(len(inner))[:len(LAX_PROBES)]]
    for f, (x, y, nx, _) in zip(probes, LAX_PROBES):
        fm.gran_center[f] = (x, y, 0.0)
        fm.gran_normal[f] = (nx, np.sqrt(1.0 - nx * nx), 0.0)
    W = random_state(rng, fm.n_cell, b=0.5)
    pot = face_flux(helio_cfg(riemann), fm, W)
    lax_seen = 0
    for f, (_, _, _, lax) in zip(probes, LAX_PROBES):
        q1 = W[fm.gran_neighbour[f, 0] - 1, :8]
        q2 = W[fm.gran_neighbour[f, 1] - 1, :8]
        want = oracle.chlld(0 if lax else riemann, fm.gran_normal[f], q1, q2)[0] * fm.gran_square[f]
        other = oracle.chlld(riemann if lax else 0, fm.gran_normal[f], q1, q2)[0] * fm.gran_square[f]
        assert not np.array_equal(want, other)  # the two solvers differ on this face
        np.testing.assert_array_equal(pot[f, :8], want)
        lax_seen += lax
    assert 0 < lax_seen < len(LAX_PROBES)


def test_v_clamp_region():
    """With the heliosphere switches the updated v of a cell is set to the
    REAL*4 0.01 exactly where the reference's tail test holds and v < 0;
    a uniform state with v < 0 is otherwise preserved (every boundary a copy
    ghost)."""
    fm = tensor_mesh(np.linspace(0.0, 1.0, 7), np.linspace(0.0, 1.0, 7))
    fm.gran_type[:] = 2
    fm.gran_type[fm.gran_neighbour[:, 1] < 0] = 1
    cells = np.arange(len(CLAMP_PROBES)) * 2
    for c, (x, y, _) in zip(cells, CLAMP_PROBES):
        fm.cell_center[c] = (x, y, 0.0)
    W = np.zeros((fm.n_cell, 16))
    W[:, 0], W[:, 1], W[:, 2], W[:, 4] = 1.0, -2.0, -0.5, 1.0
    # one cell outside the probes carries v > 0 in the tail region: untouched
    fm.cell_center[1] = (0.5, 0.03, 0.0)
    W[1, 2] = 0.25
    s = oracle.Solver(fm, W, helio_cfg(2))
    st, dt = s.face_pass(0)
    assert st == 0
    assert s.cell_pass(0, min(dt, 1e-3)) == 0
    for c, (_, _, clamped) in zip(cells, CLAMP_PROBES):
        if clamped:
            assert s.W[c, 2] == F4(0.01)
        else:
            assert s.W[c, 2] < -0.4
    assert s.W[1, 2] > 0.0
