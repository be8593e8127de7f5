"""Fused structured kernels (libsetmhd.so, structured path) against the CPU
oracle running the reference's face-list algorithm on the same grid.

Tolerances (FP64): relative L-inf <= 1e-12 per conserved variable after one
step, relative L1 <= 1e-9 after 500 steps.  The structured kernels use the
grid-axis tangent basis and shared reciprocals, so they agree with the
oracle to rounding, not bit for bit.  Mixed form: relative L1 density error
against the FP64 kernels <= 1e-5."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from set_amd import mesh, solver

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def oracle_run(g, W, nsteps, time_scheme=0):
    fm = mesh.face_mesh(g)
    s = oracle.Solver(fm, W, oracle.cartesian_config(time_scheme=time_scheme))
    s.dt_next = s.dt_scan()
    dt0 = s.dt_next
    assert s.run(nsteps) == 0
    return s, dt0


def gpu_run(g, W, nsteps, precision=0, time_scheme=0, tile=(0, 0), kchunk=0):
    with solver.Context(solver.cartesian_config(precision=precision, time_scheme=time_scheme,
                                                tile=tile, kchunk=kchunk)) as c:
        c.upload_structured(g, W)
        dt0 = c.dt_from_state()
        c.step(nsteps)
        cp, t = c.download()
    return cp, t, dt0


def cons(cp):
    return mesh.prim_to_cons(cp[:, :8])


@pytest.mark.parametrize("nsteps,bound,order", [(1, 1e-12, "inf"), (500, 1e-9, "l1")])
def test_cfg1_2d_parity(nsteps, bound, order):
    case(f"structured 2-D 400x200 HLLD Euler FP64 steps={nsteps}")
    g = mesh.grid_2d(400, 200)
    W = mesh.heliosphere_state(g)
    s, dto = oracle_run(g, W, nsteps)
    cp, t, dtg = gpu_run(g, W, nsteps)
    assert dtg == pytest.approx(dto, rel=1e-14)
    assert t == pytest.approx(s.t, rel=1e-9)
    e = rel_err(cons(cp), cons(s.W), order)
    print("rel err per var", e)
    assert np.all(e <= bound), e


@pytest.mark.parametrize("nx,kchunk", [(600, 1), (600, 2), (600, 5), (600, 64), (1100, 1), (1100, 7)])
def test_row_chunks_agree_2d(nx, kchunk):
    """2-D grids march row segments up chunks of kchunk rows: 62-column wave
    rows up to 1024 columns (600: nine full segments and one of 42), 254-column
    block rows beyond (1100: four full and one of 84).  The row below a chunk
    and the face above it take the owned rows' code, so any chunking gives
    the same bits."""
    case(f"structured 2-D {nx}x70 kchunk={kchunk} vs 16, steps=20")
    g = mesh.grid_2d(nx, 70)
    W = mesh.heliosphere_state(g)
    ref, tr, _ = gpu_run(g, W, 20, kchunk=16)
    cp, t, _ = gpu_run(g, W, 20, kchunk=kchunk)
    e = rel_err(cons(cp), cons(ref), "inf")
    print("rel Linf per var", e)
    assert t == tr
    np.testing.assert_array_equal(cp, ref)


@pytest.mark.parametrize("tile", [(16, 16), (64, 4), (16, 8)])
def test_3d_tiles_agree(tile):
    """3-D tile shapes against 32x8 (ragged tiles in x and y), FP64."""
    case(f"structured 3-D 70x36x20 tile={tile} vs 32x8, steps=10")
    g = mesh.grid_3d(70, 36, 20)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    ref, tr, _ = gpu_run(g, W, 10, tile=(32, 8))
    cp, t, _ = gpu_run(g, W, 10, tile=tile)
    e = rel_err(cons(cp), cons(ref), "inf")
    print("rel Linf per var", e, "bitwise", np.array_equal(cp, ref))
    assert t == tr
    assert np.all(e <= 1e-12), e


def test_3d_parity_one_step():
    case("structured 3-D 40x36x70 HLLD Euler FP64 steps=1")
    g = mesh.grid_3d(40, 36, 70)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    s, dto = oracle_run(g, W, 1)
    cp, t, dtg = gpu_run(g, W, 1)
    assert dtg == pytest.approx(dto, rel=1e-14)
    e = rel_err(cons(cp), cons(s.W), "inf")
    print("rel err per var", e)
    assert np.all(e <= 1e-12), e


def test_3d_parity_within_conditioning():
    """60 steps in 3-D.  The field components of this problem are ill
    conditioned: perturbing the oracle's own input by one ulp moves them by
    ~1e-8 (relative L1) after 60 steps, against ~1e-13 for rho, m and E.  The
    kernel-vs-oracle difference must stay within 10x of that 1-ulp
    sensitivity, and below 1e-10 where the problem is well conditioned."""
    case("structured 3-D 40x36x70 HLLD Euler FP64 steps=60 vs 1-ulp sensitivity")
    g = mesh.grid_3d(40, 36, 70)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    s, _ = oracle_run(g, W, 60)
    Wp = W.copy()
    Wp[:, :8] = np.nextafter(Wp[:, :8], np.inf)
    sp, _ = oracle_run(g, Wp, 60)
    sens = rel_err(cons(sp.W), cons(s.W), "l1")
    cp, t, _ = gpu_run(g, W, 60)
    e = rel_err(cons(cp), cons(s.W), "l1")
    print("rel L1 per var", e, "1-ulp sensitivity", sens)
    assert np.all(e <= np.maximum(1e-10, 10.0 * sens)), (e, sens)
    assert np.all(e[:5] <= 1e-10), e


@pytest.mark.parametrize("shape", [(32, 28, 24), (64, 64, 64)])
def test_3d_parity_500_steps(shape):
    """The north-star bound in 3-D: relative L1 <= 1e-9 after 500 steps for
    rho, m and E (their 1-ulp input sensitivity after 500 steps is ~1e-12);
    the field components, whose 1-ulp sensitivity reaches ~5e-4 by then, are
    held to 10x the larger of the oracle's sensitivities to a one-ulp step of
    its input up and down.  64^3: 262 144 cells, 16x16 tiles without ragged
    edges, eight z-chunks of 8 planes per column."""
    case(f"structured 3-D {shape} HLLD Euler FP64 steps=500 vs oracle and 1-ulp sensitivity")
    g = mesh.grid_3d(*shape)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    s, _ = oracle_run(g, W, 500)
    sens = np.zeros(8)
    sens_t = 0.0  # the time is the sum of 500 CFL steps, which see the field
    for d in (np.inf, -np.inf):
        Wp = W.copy()
        Wp[:, :8] = np.nextafter(Wp[:, :8], d)
        sp, _ = oracle_run(g, Wp, 500)
        sens = np.maximum(sens, rel_err(cons(sp.W), cons(s.W), "l1"))
        sens_t = max(sens_t, abs(sp.t - s.t) / s.t)
    cp, t, _ = gpu_run(g, W, 500)
    e = rel_err(cons(cp), cons(s.W), "l1")
    et = abs(t - s.t) / s.t
    print("rel L1 per var", e, "1-ulp sensitivity", sens, "time", et, "its sensitivity", sens_t)
    # each dt is the CFL minimum over the faces of the state it steps, formed
    # with the fast speeds' own rounding: the time agrees as the state does
    assert et <= 1e-9 and et <= max(1e-12, 10.0 * sens_t, 10.0 * e[:5].max()), (et, sens_t, e)
    assert np.all(sens[:5] < 1e-10), sens
    assert np.all(e[:5] <= 1e-9), e
    assert np.all(e[5:] <= 10.0 * sens[5:]), (e, sens)


def test_mixed_against_oracle_3d():
    """The precision form against the FP64 oracle itself in 3-D: 64^3, the
    precision config's 200 steps, per-variable relative L1 printed; density
    within 1e-5 and the momenta and energy with it (the field components are
    the ill-conditioned ones; their error is stated, not bounded)."""
    case("structured 3-D 64^3 mixed vs oracle steps=200")
    g = mesh.grid_3d(64, 64, 64)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    s, _ = oracle_run(g, W, 200)
    cp, t, _ = gpu_run(g, W, 200, precision=1)
    e = rel_err(cons(cp), cons(s.W), "l1")
    print("mixed vs oracle rel L1 per var (rho mx my mz E bx by bz)", e)
    assert np.all(e[:5] <= 1e-5), e


def test_heun_parity():
    case("structured 2-D 200x100 Heun RK2 FP64 steps=40")
    g = mesh.grid_2d(200, 100)
    W = mesh.heliosphere_state(g)
    s, _ = oracle_run(g, W, 40, time_scheme=1)
    cp, t, _ = gpu_run(g, W, 40, time_scheme=1)
    e = rel_err(cons(cp), cons(s.W), "l1")
    print("rel err per var", e)
    assert np.all(e <= 1e-10), e


def test_heun_parity_3d():
    """Heun RK2 (two launches per step, stage 2 averaging in place) in 3-D
    against the oracle: ragged 16x16 tiles, 60 steps; rho, m, E to 1e-10,
    the field components to 10x their one-ulp sensitivity."""
    case("structured 3-D 50x42x36 Heun RK2 FP64 steps=60")
    g = mesh.grid_3d(50, 42, 36)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    s, _ = oracle_run(g, W, 60, time_scheme=1)
    Wp = W.copy()
    Wp[:, :8] = np.nextafter(Wp[:, :8], np.inf)
    sp, _ = oracle_run(g, Wp, 60, time_scheme=1)
    sens = rel_err(cons(sp.W), cons(s.W), "l1")
    cp, t, _ = gpu_run(g, W, 60, time_scheme=1)
    e = rel_err(cons(cp), cons(s.W), "l1")
    print("rel L1 per var", e, "1-ulp sensitivity", sens)
    assert t == pytest.approx(s.t, rel=1e-11)
    assert np.all(e[:5] <= 1e-10), e
    assert np.all(e[5:] <= np.maximum(10.0 * sens[5:], 1e-10)), (e, sens)


def test_mixed_precision_error():
    case("structured 2-D 400x200 mixed vs FP64 steps=500")
    g = mesh.grid_2d(400, 200)
    W = mesh.heliosphere_state(g)
    ref, _, _ = gpu_run(g, W, 500)
    cp, _, _ = gpu_run(g, W, 500, precision=1)
    e = rel_err(cons(cp), cons(ref), "l1")
    print("mixed rel L1 per var", e)
    assert e[0] <= 1e-5, e


def dropin_run(g, W, nsteps, dt0):
    ocfg = oracle.cartesian_config()
    c = solver.Config()
    c.riemann, c.recon, c.time_scheme, c.flags = ocfg.riemann, ocfg.recon, ocfg.time_scheme, ocfg.flags
    c.cfl, c.dt_cap, c.r_skip_face, c.r_hold_cell = ocfg.cfl, ocfg.dt_cap, ocfg.r_skip_face, ocfg.r_hold_cell
    with solver.Context(c) as ctx:
        ctx.upload_mesh(mesh.face_mesh(g), W)
        ctx.set_dt(dt0)
        ctx.step(nsteps)
        return ctx.download()


@pytest.mark.parametrize("shape,nsteps,order", [((400, 200, 1), 1, "inf"), ((400, 200, 1), 100, "l1"),
                                                ((48, 40, 36), 1, "inf"), ((48, 40, 36), 30, "l1")])
def test_structured_equals_dropin_on_the_same_grid(shape, nsteps, order):
    """The fused structured kernel against the drop-in face-list kernels
    (the reference's face pass + cell pass restated) on the same grid, both
    on the GPU, same dt sequence: relative L-inf <= 1e-12 after one step;
    after more steps relative L1 <= 1e-9 in 2-D, and in 3-D within 10x of
    the drop-in's own 1-ulp input sensitivity (the field components are ill
    conditioned there, see test_3d_parity_within_conditioning) and <= 1e-10
    for rho, m and E."""
    case(f"structured vs drop-in {shape} steps={nsteps}")
    nx, ny, nz = shape
    g = mesh.grid_2d(nx, ny) if nz == 1 else mesh.grid_3d(nx, ny, nz)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.1 if nz == 1 else 0.3))
    cp_s, t_s, dt0 = gpu_run(g, W, nsteps)
    cp_u, t_u = dropin_run(g, W, nsteps, dt0)
    assert t_s == pytest.approx(t_u, rel=1e-14)
    e = rel_err(cons(cp_s), cons(cp_u), order)
    print(f"rel {order} per var", e)
    if nsteps == 1:
        assert np.all(e <= 1e-12), e
    elif nz == 1:
        assert np.all(e <= 1e-9), e
    else:
        Wp = W.copy()
        Wp[:, :8] = np.nextafter(Wp[:, :8], np.inf)
        cp_p, _ = dropin_run(g, Wp, nsteps, dt0)
        sens = rel_err(cons(cp_p), cons(cp_u), order)
        print("1-ulp sensitivity", sens)
        assert np.all(e <= np.maximum(1e-10, 10.0 * sens)), (e, sens)
        assert np.all(e[:5] <= 1e-10), e


# 2-D wave rows own 62 columns per segment: 62, 63 and 125 columns end a
# segment exactly, leave one column to a second segment, and to a third;
# 254, 255 and 509 the same in 3-4 segments; 1100 columns march 254-column
# block rows (four full segments and one of 84)
TINY = [(1, 1, 1), (3, 1, 1), (1, 5, 1), (2, 2, 2), (1, 1, 7), (2, 3, 4), (33, 9, 1), (62, 3, 1),
        (63, 2, 1), (125, 2, 1), (254, 3, 1), (255, 2, 1), (509, 2, 1), (1100, 3, 1)]


def vec_err(a, b):
    """max |a-b| per conserved variable, scaled by the largest magnitude of
    its quantity (rho; momentum vector; E; field vector): components that
    stay at round-off in a small uniform box are judged against the vector
    they belong to, not against their own noise."""
    d = np.abs(a - b).max(0)
    m = np.abs(b).max(0)
    scale = np.array([m[0], *[m[1:4].max()] * 3, m[4], *[m[5:8].max()] * 3])
    return d / np.where(scale > 0, scale, 1.0)


@pytest.mark.parametrize("shape", TINY)
def test_tiny_and_ragged_grids(shape):
    """Grids narrower than a tile in every direction (one cell wide, one
    tile plus one cell) and 2-D grids at row-segment boundaries: FP64
    against the oracle after 3 steps, and the mixed form runs."""
    case(f"structured tiny grid {shape}")
    nx, ny, nz = shape
    g = mesh.grid_2d(nx, ny) if nz == 1 else mesh.grid_3d(nx, ny, nz)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    s, _ = oracle_run(g, W, 3)
    cp, _, _ = gpu_run(g, W, 3)
    e = vec_err(cons(cp), cons(s.W))
    print("scaled Linf per var", e)
    assert np.all(e <= 1e-12), e
    cpm, _, _ = gpu_run(g, W, 3, precision=1)
    assert np.all(np.isfinite(cpm))


def test_tiny_grids_in_one_batch():
    case("batch of tiny and ragged grids")
    gs = []
    for nx, ny, nz in TINY:
        g = mesh.grid_2d(nx, ny) if nz == 1 else mesh.grid_3d(nx, ny, nz)
        gs.append((g, mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))))
    with solver.Context(solver.cartesian_config()) as c:
        slots = [c.batch_add(g, W) for g, W in gs]
        c.batch_step(3)
        got = [c.batch_download(sl)[0] for sl in slots]
    for (g, W), cp in zip(gs, got):
        ref, _, _ = gpu_run(g, W, 3)
        e = vec_err(cons(cp), cons(ref))
        assert np.all(e <= 1e-12), (g, e)


@pytest.mark.parametrize("nx", [66, 65])
def test_mixed_plane_dma_matches_register_path(nx):
    """The mixed 3-D kernel with 32x8 tiles stages each plane through LDS
    DMA when nx is even (65: the register path).  Ragged tiles in x and y;
    compared with the 16x16 kernel, which always loads through registers."""
    case(f"structured 3-D {nx}x36x40 mixed, 32x8 vs 16x16, steps=5")
    g = mesh.grid_3d(nx, 36, 40)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    a, ta, _ = gpu_run(g, W, 5, precision=1, tile=(32, 8))
    b, tb, _ = gpu_run(g, W, 5, precision=1, tile=(16, 16))
    e = rel_err(cons(a), cons(b), "inf")
    print("rel Linf per var", e, "bitwise", np.array_equal(a, b))
    assert ta == tb
    assert np.all(e <= 1e-9), e


def test_mixed_heun_close_to_fp64_heun():
    case("structured 3-D 48x40x36 mixed Heun vs FP64 Heun steps=20")
    g = mesh.grid_3d(48, 40, 36)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    ref, tr, _ = gpu_run(g, W, 20, time_scheme=1)
    cp, tm, _ = gpu_run(g, W, 20, precision=1, time_scheme=1)
    e = rel_err(cons(cp), cons(ref), "l1")
    print("mixed Heun rel L1 per var", e)
    assert e[0] <= 1e-5, e
    assert tm == pytest.approx(tr, rel=1e-5)


def test_contexts_in_two_threads_are_independent():
    """Two contexts stepping at the same time from two host threads (each
    owns its stream and device state) give the results each gives alone."""
    import threading
    case("two contexts in two threads")
    g1, g2 = mesh.grid_3d(40, 36, 30), mesh.grid_2d(300, 150)
    W1 = mesh.heliosphere_state(g1, wind=mesh.Wind(r_in=0.3))
    W2 = mesh.heliosphere_state(g2)
    ref1, _, _ = gpu_run(g1, W1, 40)
    ref2, _, _ = gpu_run(g2, W2, 40, precision=1)
    out = {}

    def work(key, g, W, precision):
        out[key] = gpu_run(g, W, 40, precision=precision)[0]

    ts = [threading.Thread(target=work, args=(1, g1, W1, 0)),
          threading.Thread(target=work, args=(2, g2, W2, 1))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    np.testing.assert_array_equal(out[1], ref1)
    np.testing.assert_array_equal(out[2], ref2)


def test_free_stream_256cube_exact():
    """Full size: a uniform state is a fixed point, bit for bit (every face
    sees the same two states, so every face sum cancels exactly)."""
    case("structured 3-D 256^3 free stream")
    g = mesh.grid_3d(256, 256, 256)
    W = np.zeros((g.n_cell, 16))
    W[:, :8] = [0.1, -2.5, 0.3, -0.2, 0.1, 0.26, 0.15, 0.05]
    with solver.Context(solver.cartesian_config(r_hold=-1.0)) as c:
        c.upload_structured(g, W)
        c.dt_from_state()
        c.step(3)
        cp, _ = c.download()
    np.testing.assert_array_equal(cons(cp), cons(W))


def _phys_flux(U, axis):
    """Ideal-MHD physical flux of conserved states U (N x 8: rho, m, E, b with
    b = B/sqrt(4 pi) and E carrying |b|^2/2) along grid axis `axis`."""
    rho, m, E, b = U[:, 0], U[:, 1:4], U[:, 4], U[:, 5:8]
    u = m / rho[:, None]
    bb = (b * b).sum(1)
    pt = (mesh.GAMMA - 1.0) * (E - 0.5 * (m * u).sum(1) - 0.5 * bb) + 0.5 * bb
    un, bn = u[:, axis], b[:, axis]
    F = np.empty_like(U)
    F[:, 0] = m[:, axis]
    F[:, 1:4] = m * un[:, None] - bn[:, None] * b
    F[:, 1 + axis] += pt
    F[:, 4] = (E + pt) * un - bn * (u * b).sum(1)
    F[:, 5:8] = un[:, None] * b - bn[:, None] * u
    return F


@pytest.mark.parametrize("precision", [0, 1])
def test_conservation_256cube(precision):
    """Full size: one forward-Euler step changes the grid's total of every
    conserved variable by exactly what crosses the boundary.  With no held
    cells, sum_i dU_i = -dt/h sum_faces F(U_b).n_out over the six sides, the
    outflow faces seeing the boundary cell on both sides (HLLD of equal
    states is the physical flux).  FP64: to 1e-12 of the grid's total |U|
    (a dropped or mis-signed face flux moves it by ~1e-9); mixed form: the
    interior still cancels in FP64, the boundary fluxes come from FP32 HLLD,
    so 1e-6 of the gross boundary flux is added."""
    case(f"structured 3-D 256^3 conservation, precision {precision}")
    n = 256
    g = mesh.grid_3d(n, n, n)
    W = mesh.heliosphere_state(g)
    with solver.Context(solver.cartesian_config(r_hold=-1.0, precision=precision)) as c:
        c.upload_structured(g, W)
        c.dt_from_state()
        c.step(1)
        cp, dt = c.download()
    U0 = cons(W).reshape(n, n, n, 8)  # [k][j][i]
    U1 = cons(cp).reshape(n, n, n, 8)
    lhs = (U1 - U0).reshape(-1, 8).sum(0)
    h = 2.0 / n
    net = np.zeros(8)
    gross = np.zeros(8)
    for axis, lo, hi in ((0, U0[:, :, 0], U0[:, :, -1]), (1, U0[:, 0], U0[:, -1]), (2, U0[0], U0[-1])):
        Fl = _phys_flux(lo.reshape(-1, 8), axis)
        Fh = _phys_flux(hi.reshape(-1, 8), axis)
        net += Fh.sum(0) - Fl.sum(0)
        gross += np.abs(Fh).sum(0) + np.abs(Fl).sum(0)
    rhs = -dt / h * net
    scale = np.abs(U0).reshape(-1, 8).sum(0)
    err = np.abs(lhs - rhs)
    tol = 1e-12 * scale + (0.0 if precision == 0 else 1e-6 * dt / h * gross)
    print("sum dU", lhs, "\n-dt/h boundary flux", rhs, "\nerr", err, "\ntol", tol)
    assert np.all(err <= tol), (err, tol)  # bz stays 0 exactly: B has no z part, tol 0


def test_mirror_symmetry_256cube_hydro():
    """Full size: with B = 0 the heliosphere set-up (no density noise) is
    symmetric under z -> -z and the scheme keeps it to rounding.  (With
    B != 0 the reference's HLLC/HLLD is not mirror symmetric: on a face
    where the normal field jumps, chlld builds the star states with the
    HLL-averaged Bn while the outer fluxes keep their own Bn, so odd flux
    components do not cancel; the oracle shows the same.)"""
    case("structured 3-D 256^3 z-mirror symmetry, B = 0")
    g = mesh.grid_3d(256, 256, 256)
    W = mesh.heliosphere_state(g, inflow=mesh.Inflow(noise=0.0, b=0.0))
    with solver.Context(solver.cartesian_config()) as c:
        c.upload_structured(g, W)
        c.dt_from_state()
        c.step(6)
        cp, _ = c.download()
    U = cons(cp).reshape(256, 256, 256, 8)
    M = U[::-1].copy()
    M[..., 3] *= -1.0  # z momentum
    scale = np.abs(U).reshape(-1, 8).max(0)
    err = np.abs(U - M).reshape(-1, 8).max(0) / np.where(scale > 0, scale, 1.0)
    print("mirror err per var", err)
    assert np.all(err <= 1e-13), err


@pytest.mark.parametrize("precision,kchunk", [(0, 1), (0, 7), (0, 60), (1, 1), (1, 7), (1, 60)])
def test_zchunk_invariance_3d(precision, kchunk):
    """The plane below a z-chunk and the face above it take the owned planes'
    code, so any chunking gives the same bits (70x36x60: ragged 32x8 tiles;
    the mixed form stages planes by DMA)."""
    case(f"structured 3-D 70x36x60 precision={precision} kchunk={kchunk} vs 32, steps=10")
    g = mesh.grid_3d(70, 36, 60)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    out = []
    for kc in (32, kchunk):
        with solver.Context(solver.cartesian_config(precision=precision, kchunk=kc)) as c:
            c.upload_structured(g, W)
            c.dt_from_state()
            c.step(10)
            out.append(c.download()[0])
    np.testing.assert_array_equal(out[1], out[0])


@pytest.mark.parametrize("kchunk", [256, 17])
def test_zchunk_invariance_256cube(kchunk):
    """Full size: results do not depend on the z-chunk length (bitwise)."""
    case(f"structured 3-D 256^3 kchunk={kchunk} vs 32")
    g = mesh.grid_3d(256, 256, 256)
    W = mesh.heliosphere_state(g)
    out = []
    for kc in (32, kchunk):
        with solver.Context(solver.cartesian_config(kchunk=kc)) as c:
            c.upload_structured(g, W)
            c.dt_from_state()
            c.step(3)
            out.append(cons(c.download()[0]))
    np.testing.assert_array_equal(out[1], out[0])


def test_mixed_vs_fp64_256cube():
    """Full size, the precision form's bound: relative L1 density error of
    the mixed form against FP64 after 200 steps <= 1e-5."""
    case("structured 3-D 256^3 mixed vs FP64, 200 steps")
    g = mesh.grid_3d(256, 256, 256)
    W = mesh.heliosphere_state(g)
    res = []
    for prec in (0, 1):
        with solver.Context(solver.cartesian_config(precision=prec)) as c:
            c.upload_structured(g, W)
            c.dt_from_state()
            c.step(200)
            res.append(cons(c.download()[0]))
    e = rel_err(res[1], res[0], "l1")
    print("mixed vs fp64 rel L1 per var", e)
    assert e[0] <= 1e-5, e
