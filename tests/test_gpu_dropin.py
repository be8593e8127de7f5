"""Drop-in face-list kernels (libsetmhd.so, unstructured path) against the
CPU oracle on the same seeded inputs.  Both restate the reference in the
same operation order with contraction off, so results agree bit for bit
except where a transcendental (log in Source_MK) is evaluated by different
math libraries."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from set_amd import mesh, solver, workloads

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def gpu_cfg(o):
    c = solver.Config()
    c.riemann, c.recon, c.time_scheme, c.flags = o.riemann, o.recon, o.time_scheme, o.flags
    c.cfl, c.dt_cap, c.r_skip_face, c.r_hold_cell = o.cfl, o.dt_cap, o.r_skip_face, o.r_hold_cell
    return c


def run_both(fm, W, ocfg, nsteps, dt0):
    s = oracle.Solver(fm, W, ocfg)
    s.dt_next = dt0
    assert s.run(nsteps) == 0
    with solver.Context(gpu_cfg(ocfg)) as c:
        c.upload_mesh(fm, W)
        c.set_dt(dt0)
        c.step(nsteps)
        cp, t = c.download()
        _, dtn = c.sync()
        pot = c.download_potok()
    return s, cp, t, dtn, pot


@pytest.mark.parametrize("nsteps", [1, 25])
def test_cartesian_hlld_bitwise(nsteps):
    case(f"dropin cartesian 2-D 100x50 HLLD Euler steps={nsteps}")
    g = mesh.grid_2d(100, 50)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g)
    cfg = oracle.cartesian_config()
    dt0 = oracle.Solver(fm, W, cfg).dt_scan()
    s, cp, t, dtn, pot = run_both(fm, W, cfg, nsteps, dt0)
    np.testing.assert_array_equal(cp[:, :8], s.W[:, :8])
    np.testing.assert_array_equal(cp[:, 8:], W[:, 8:])
    np.testing.assert_array_equal(pot, s.potok)
    assert t == s.t and dtn == s.dt_next


@pytest.mark.parametrize("riemann", [0, 1, 2, 3])
@pytest.mark.parametrize("recon,scheme", [(0, 0), (1, 1)])
def test_all_solvers_bitwise(riemann, recon, scheme):
    case(f"dropin riemann={riemann} recon={recon} scheme={scheme}")
    g = mesh.grid_2d(40, 20)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.2, speed=3.0))
    cfg = oracle.cartesian_config(riemann=riemann, recon=recon, time_scheme=scheme)
    dt0 = oracle.Solver(fm, W, cfg).dt_scan()
    s, cp, t, dtn, _ = run_both(fm, W, cfg, 6, dt0)
    np.testing.assert_array_equal(cp[:, :8], s.W[:, :8])
    assert dtn == s.dt_next


def test_3d_bitwise():
    case("dropin 3-D 12x10x8 HLLD Euler")
    g = mesh.grid_3d(12, 10, 8)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.35, speed=3.0))
    cfg = oracle.cartesian_config()
    dt0 = oracle.Solver(fm, W, cfg).dt_scan()
    s, cp, t, dtn, _ = run_both(fm, W, cfg, 8, dt0)
    np.testing.assert_array_equal(cp[:, :8], s.W[:, :8])


helio_mesh = workloads.heliosphere_face_mesh  # half-plane heliosphere, y = 0 the axis


def test_reference_semantics():
    """HLLC + LAX wedge, minmod TVD, Heun RK2, axisymmetric and charge-
    exchange sources, w = B = 0, clamps; reference initial dt 2e-9."""
    case("dropin reference semantics 2-D 120x60 RK2")
    fm, W = helio_mesh(120, 60)
    cfg = oracle.reference_config()
    s, cp, t, dtn, _ = run_both(fm, W, cfg, 12, 2.0e-9)
    err = np.abs(cp[:, :8] - s.W[:, :8]).max(0) / np.maximum(np.abs(s.W[:, :8]).max(0), 1e-300)
    assert np.all(err <= 1e-12), err
    assert t == pytest.approx(s.t, rel=1e-13)


@pytest.mark.parametrize("which", ["reference", "cartesian3d"])
def test_fast_build_agrees_with_oracle(which):
    """SETMHD_F_FAST: the same kernel text with FMA contraction and
    reciprocal-based division.  Not bit for bit; relative L-inf <= 1e-12 per
    variable against the oracle after several steps."""
    case(f"dropin fast build {which}")
    if which == "reference":
        fm, W = helio_mesh(120, 60)
        ocfg, nsteps, dt0 = oracle.reference_config(), 12, 2.0e-9
    else:
        g = mesh.grid_3d(12, 10, 8)
        fm = mesh.face_mesh(g)
        W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.35, speed=3.0))
        ocfg, nsteps = oracle.cartesian_config(), 8
        dt0 = oracle.Solver(fm, W, ocfg).dt_scan()
    s = oracle.Solver(fm, W, ocfg)
    s.dt_next = dt0
    assert s.run(nsteps) == 0
    cfg = gpu_cfg(ocfg)
    cfg.flags |= solver.F_FAST
    with solver.Context(cfg) as c:
        c.upload_mesh(fm, W)
        c.set_dt(dt0)
        c.step(nsteps)
        cp, t = c.download()
    err = np.abs(cp[:, :8] - s.W[:, :8]).max(0) / np.maximum(np.abs(s.W[:, :8]).max(0), 1e-300)
    print("rel Linf per var", err)
    assert np.all(err <= 1e-12), err
    assert t == pytest.approx(s.t, rel=1e-13)


@pytest.mark.parametrize("nsteps,bound,order", [(1, 1e-12, "inf"), (500, 1e-9, "l1")])
def test_fast_build_hlld_plain_2d(nsteps, bound, order):
    """SETMHD_F_FAST with HLLD and B != 0 (chlld with FMA contraction and
    reciprocal-based division) on the plain 400x200 set-up against the
    oracle: relative L-inf <= 1e-12 after one step, relative L1 <= 1e-9 on
    all 8 conserved variables after 500, the bounds of the exact path."""
    case(f"dropin fast build HLLD 400x200 steps={nsteps}")
    g = mesh.grid_2d(400, 200)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g)
    ocfg = oracle.cartesian_config()
    dt0 = oracle.Solver(fm, W, ocfg).dt_scan()
    s = oracle.Solver(fm, W, ocfg)
    s.dt_next = dt0
    assert s.run(nsteps) == 0
    cfg = gpu_cfg(ocfg)
    cfg.flags |= solver.F_FAST
    with solver.Context(cfg) as c:
        c.upload_mesh(fm, W)
        c.set_dt(dt0)
        c.step(nsteps)
        cp, t = c.download()
    e = rel_err(mesh.prim_to_cons(cp[:, :8]), mesh.prim_to_cons(s.W[:, :8]), order)
    print("rel err per var", order, e)
    assert np.all(e <= bound), e
    assert t == pytest.approx(s.t, rel=1e-12)


@pytest.mark.parametrize("which", ["reference", "cartesian3d"])
def test_reorder_bitwise(which):
    """SETMHD_F_REORDER on a randomly renumbered mesh: the device holds its
    own (Morton) numbering while every call takes and returns the caller's.
    State, snapshot, face fluxes, probes, time and next dt equal bit for bit
    those of the original mesh run without reordering, mapped through the
    renumbering; so does the renumbered mesh run without reordering."""
    if which == "reference":
        fm, W = helio_mesh(120, 60)
        ocfg, nsteps, dt0 = oracle.reference_config(), 12, 2.0e-9
    else:
        g = mesh.grid_3d(24, 20, 16)
        fm = mesh.face_mesh(g)
        W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
        ocfg, nsteps, dt0 = oracle.cartesian_config(), 10, None
    case(f"dropin reorder {which} {fm.n_cell} cells, {nsteps} steps")
    rng = np.random.default_rng(4)
    cp, gp = rng.permutation(fm.n_cell), rng.permutation(fm.n_gran)
    fr, Wr = mesh.renumber_face_mesh(fm, cp, gp), W[cp]
    probe = np.array([1, fm.n_cell // 3, fm.n_cell])  # numbers in the renumbered mesh
    out = {}
    for name, m, w, flags, ids in (("base", fm, W, 0, cp[probe - 1] + 1),
                                   ("renumbered", fr, Wr, 0, probe),
                                   ("reordered", fr, Wr, solver.F_REORDER, probe)):
        cfg = gpu_cfg(ocfg)
        cfg.flags |= flags
        with solver.Context(cfg) as c:
            c.upload_mesh(m, w)
            if dt0 is None:
                dt0 = c.dt_from_state()
            c.set_dt(dt0)
            c.step(nsteps)
            c.download_async()
            cpar, t = c.download()
            snap, ts = c.download_wait()
            _, dtn = c.sync()
            out[name] = dict(cp=cpar, snap=snap, t=t, ts=ts, dtn=dtn, pot=c.download_potok(),
                             probe=c.probe(ids)[0])
    b, r, o = out["base"], out["renumbered"], out["reordered"]
    np.testing.assert_array_equal(r["cp"], b["cp"][cp])
    np.testing.assert_array_equal(r["pot"], b["pot"][gp])
    for k in ("cp", "snap", "pot", "probe"):
        np.testing.assert_array_equal(o[k], r[k])
    np.testing.assert_array_equal(o["snap"], o["cp"])
    np.testing.assert_array_equal(o["probe"], b["probe"])
    np.testing.assert_array_equal(o["probe"], o["cp"][probe - 1, :8])
    assert o["t"] == r["t"] == b["t"] == o["ts"]
    assert o["dtn"] == r["dtn"] == b["dtn"]


def retyped_helio_mesh(nx, ny):
    """The half-plane heliosphere mesh with its outer sides retyped: x = -1 a
    reflecting wall (Gran_type 2 with no second cell, what SET writes for a
    face whose other cell is not exported, Setka.cpp:1456-1483), y = 1 type
    1 and x = 1 type 3 without a second cell (copy ghosts); y = 0 stays the
    axis (type 5)."""
    fm, W = helio_mesh(nx, ny)
    out = fm.gran_neighbour[:, 1] < 0
    n = fm.gran_normal
    fm.gran_type[out & (n[:, 0] == -1.0)] = mesh.TYPE_INTERIOR
    fm.gran_type[out & (n[:, 1] == 1.0)] = 1
    fm.gran_type[out & (n[:, 0] == 1.0)] = 3
    return fm, W


@pytest.mark.parametrize("which", ["reference", "cartesian"])
def test_wall_and_copy_ghosts(which):
    """Every boundary type on the GPU against the oracle: the wall, types 1
    and 3 and the axis, under SET's semantics (TVD, Heun, sources; <= 1e-12
    for Source_MK's log) and under plain first-order HLLD with Euler (bit
    for bit, fluxes and next dt included)."""
    fm, W = retyped_helio_mesh(120, 60)
    assert set(np.unique(fm.gran_type)) == {1, 2, 3, 5}
    if which == "reference":
        ocfg, nsteps, dt0 = oracle.reference_config(), 12, 2.0e-9
    else:
        ocfg, nsteps = oracle.cartesian_config(riemann=3, recon=1, time_scheme=1, r_hold=0.1), 40
        dt0 = oracle.Solver(fm, W, ocfg).dt_scan()
    case(f"dropin wall + types 1/3/5 {which} 120x60 steps={nsteps}")
    s, cp, t, dtn, pot = run_both(fm, W, ocfg, nsteps, dt0)
    if which == "reference":
        err = np.abs(cp[:, :8] - s.W[:, :8]).max(0) / np.maximum(np.abs(s.W[:, :8]).max(0), 1e-300)
        assert np.all(err <= 1e-12), err
        assert t == pytest.approx(s.t, rel=1e-13)
    else:
        np.testing.assert_array_equal(cp[:, :8], s.W[:, :8])
        np.testing.assert_array_equal(pot, s.potok)
        assert t == s.t and dtn == s.dt_next


def test_density_failure_reported():
    case("dropin forced rho<0 -> SETMHD_ERR_RHO")
    g = mesh.grid_2d(40, 20)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g)
    with solver.Context(gpu_cfg(oracle.cartesian_config())) as c:
        c.upload_mesh(fm, W)
        c.set_dt(10.0)  # far beyond CFL: densities go negative
        with pytest.raises(solver.SetMHDError) as e:
            c.step(1)
        assert e.value.status == solver.ERR_RHO


def test_upload_validation():
    case("dropin Set_Storage validation")
    g = mesh.grid_2d(8, 4)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g)
    bad = mesh.face_mesh(g)
    bad.gran_neighbour = bad.gran_neighbour.copy()
    bad.gran_neighbour[3, 0] = fm.n_cell + 1
    with solver.Context(gpu_cfg(oracle.cartesian_config())) as c:
        with pytest.raises(solver.SetMHDError) as e:
            c.upload_mesh(bad, W)
        assert e.value.status == solver.ERR_ARG
        W2 = W.copy()
        W2[5, 0] = 0.0
        with pytest.raises(solver.SetMHDError):
            c.upload_mesh(fm, W2)
        # Set_Storage compares rho with the default-REAL literal 1e-9
        # (9.99999972e-10): a density between the two is accepted, as the
        # Fortran and Python readers accept it
        W2[5, 0] = 9.999999985e-10
        c.upload_mesh(fm, W2)
        for col in (8, 11, 12, 15):  # Cell_par(9), (12), (13), (16) must be >= 0
            W3 = W.copy()
            W3[7, col] = -1e-3
            with pytest.raises(solver.SetMHDError) as e:
                c.upload_mesh(fm, W3)
            assert e.value.status == solver.ERR_ARG


@pytest.mark.parametrize("mode", ["reference", "cartesian"])
def test_fortran_driver(tmp_path, mode):
    """The Fortran entry point (set_amd/fortran/fcmhd_driver, iso_c_binding
    to the C ABI) reads an FCMHD input file, runs the steps on the GPU and
    writes the Save_Storage file; compared with the oracle from dt = 2e-9."""
    import os
    import subprocess

    from set_amd import fcmhd
    case(f"fortran driver {mode}")
    exe = os.path.join(os.path.dirname(solver.HERE), "set_amd", "fortran", "fcmhd_driver")
    if mode == "reference":
        fm, W = helio_mesh(60, 30)
        cfg = oracle.reference_config()
    else:
        g = mesh.grid_2d(60, 30)
        fm = mesh.face_mesh(g)
        W = mesh.heliosphere_state(g)
        cfg = oracle.cartesian_config()
    fin, fout = tmp_path / "FCMHD_3.bin", tmp_path / "FCMHD_3.5_out.bin"
    fcmhd.write_input(fin, fm, W)
    args = [exe, str(fin), str(fout), "12", "5", ""] + (["cartesian"] if mode == "cartesian" else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    t, cp = fcmhd.fill_data(fout, fm.n_cell)
    s = oracle.Solver(fm, W, cfg)
    s.dt_next = 2.0e-9
    assert s.run(12) == 0
    if mode == "cartesian":
        np.testing.assert_array_equal(cp[:, :8], s.W[:, :8])
    else:
        err = np.abs(cp[:, :8] - s.W[:, :8]).max(0) / np.maximum(np.abs(s.W[:, :8]).max(0), 1e-300)
        assert np.all(err <= 1e-12), err
    np.testing.assert_array_equal(cp[:, 8:], W[:, 8:])
    assert t == pytest.approx(s.t, rel=1e-13)


def test_fortran_driver_outputs(tmp_path):
    """The driver's observability outputs: the probe series (setmhd_probe,
    write_T_rho layout) and the Tecplot dumps (Print_mult and Print layouts)
    hold the run's state; the dumps equal set_amd.tecplot's text."""
    import os
    import subprocess

    from set_amd import fcmhd, tecplot
    case("fortran driver probes and dumps")
    exe = os.path.join(os.path.dirname(solver.HERE), "set_amd", "fortran", "fcmhd_driver")
    fm, W = helio_mesh(60, 30)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    fcmhd.write_input(fin, fm, W)
    cells = [1, 100, 1800]
    (tmp_path / "probes.txt").write_text("".join(f"{c}\n" for c in cells))
    r = subprocess.run([exe, str(fin), str(fout), "12", "5", "-", "reference", "probes.txt", "dump"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    t, cp = fcmhd.fill_data(fout, fm.n_cell)
    series = (tmp_path / "1D_rhoT.txt").read_text().splitlines()
    assert len(series) == 3  # after steps 5, 10 and 12
    last = [float(v) for v in series[-1].split()]
    np.testing.assert_allclose(last, [t * tecplot.SCALE_T] + list(cp[np.array(cells) - 1, 0] * tecplot.SCALE_RHO),
                               rtol=1e-14)
    tecplot.write_density(tmp_path / "py_density.txt", fm.cell_center, cp, t)
    tecplot.write_state(tmp_path / "py_state.txt", fm.cell_center, cp)
    for f_out, py in (("output_data_3.txt", "py_density.txt"), ("output_data.txt", "py_state.txt")):
        a = [l.rstrip() for l in (tmp_path / f_out).read_text().splitlines()]
        b = [l.rstrip() for l in (tmp_path / py).read_text().splitlines()]
        assert a == b, (f_out, next((x, y) for x, y in zip(a, b) if x != y) if len(a) == len(b) else len(a))
    # the same run with the mesh renumbered on the device: every output file
    # (state, probe series, dumps) byte for byte the same
    rdir = tmp_path / "reorder"
    rdir.mkdir()
    (rdir / "probes.txt").write_text((tmp_path / "probes.txt").read_text())
    r = subprocess.run([exe, str(fin), str(rdir / "out.bin"), "12", "5", "-", "reference", "probes.txt", "dump",
                        "reorder"], capture_output=True, text=True, timeout=120, cwd=rdir)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    for name in ("out.bin", "1D_rhoT.txt", "output_data_3.txt", "output_data.txt"):
        assert (rdir / name).read_bytes() == (tmp_path / name).read_bytes(), name


@pytest.mark.parametrize("mode", ["cartesian", "reference"])
def test_fortran_driver_restart(tmp_path, mode):
    """Restart through the Fortran driver (Save_Storage / Fill_data,
    Storage_module.f90:163-209; FCMHD.f90:129,248).  N steps, save, restart
    from the saved file:
      * with 'continue' (the saved lagged dt is taken up) M more steps give
        the file of N+M steps straight, byte for byte;
      * as the reference restarts (dt back to 2e-9, cuf_kernel.cuf:758) the
        run equals the oracle started from the saved state and time.
    The time is kept on resume (the reference resets it, cuf_kernel.cuf:353)."""
    import os
    import subprocess

    from set_amd import fcmhd
    case(f"fortran driver restart {mode}")
    exe = os.path.join(os.path.dirname(solver.HERE), "set_amd", "fortran", "fcmhd_driver")
    if mode == "reference":
        fm, W = helio_mesh(60, 30)
        cfg = oracle.reference_config()
    else:
        g = mesh.grid_2d(60, 30)
        fm = mesh.face_mesh(g)
        W = mesh.heliosphere_state(g)
        cfg = oracle.cartesian_config()
    fin = tmp_path / "in.bin"
    fcmhd.write_input(fin, fm, W)
    kind = ["cartesian"] if mode == "cartesian" else ["reference"]

    def drive(out, n, restart="-", *extra):
        r = subprocess.run([exe, str(fin), str(tmp_path / out), str(n), "5", restart] + kind +
                           ["-", "-", "-"] + list(extra), capture_output=True, text=True, timeout=120)
        print(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
        return fcmhd.fill_data(tmp_path / out, fm.n_cell)

    t_full, cp_full = drive("full.bin", 17)
    t1, cp1 = drive("part1.bin", 10)
    t2, cp2 = drive("part2.bin", 7, str(tmp_path / "part1.bin"), "continue")
    assert (tmp_path / "part2.bin").read_bytes() == (tmp_path / "full.bin").read_bytes()
    assert (tmp_path / "part2.bin.dt").read_bytes() == (tmp_path / "full.bin.dt").read_bytes()
    assert t2 == t_full and t1 < t2
    # the reference's own restart: time kept, dt back to 2e-9
    t3, cp3 = drive("part3.bin", 7, str(tmp_path / "part1.bin"))
    s = oracle.Solver(fm, cp1, cfg)
    s.t, s.dt_next = t1, 2.0e-9
    assert s.run(7) == 0
    if mode == "cartesian":
        np.testing.assert_array_equal(cp3, s.W)
        assert t3 == s.t
    else:
        err = np.abs(cp3[:, :8] - s.W[:, :8]).max(0) / np.maximum(np.abs(s.W[:, :8]).max(0), 1e-300)
        assert np.all(err <= 1e-12), err
        assert t3 == pytest.approx(s.t, rel=1e-13)
        np.testing.assert_array_equal(cp3[:, 8:], cp1[:, 8:])


def run_gpu(fm, W, cfg, nsteps, dt0, flags=0):
    cfg.flags |= flags
    with solver.Context(cfg) as c:
        c.upload_mesh(fm, W)
        c.set_dt(dt0)
        c.step(nsteps)
        cp, t = c.download()
        _, dtn = c.sync()
        pot = c.download_potok()
    return cp, t, dtn, pot


def assert_same_run(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[3], b[3])
    assert a[1] == b[1] and a[2] == b[2]


@pytest.mark.parametrize("build", ["exact", "fast", "reorder"])
def test_per_cell_heun_same_bits(build):
    """SETMHD_F_PER_CELL (one launch per Heun stage, each cell forming its
    own faces) gives the bits of the face pass + cell pass: state, Gran_POTOK,
    time and next dt, at the reference's semantics (HLLC + LAX wedge, TVD,
    axisymmetric and charge-exchange sources, held cells)."""
    case(f"dropin per-cell Heun vs two launches per stage, {build}, heliosphere 120x60")
    fm, W = helio_mesh(120, 60)
    extra = {"exact": 0, "fast": solver.F_FAST, "reorder": solver.F_REORDER}[build]
    ocfg = oracle.reference_config()
    a = run_gpu(fm, W, gpu_cfg(ocfg), 12, 2.0e-9, extra)
    b = run_gpu(fm, W, gpu_cfg(ocfg), 12, 2.0e-9, extra | solver.F_PER_CELL)
    assert_same_run(a, b)


def test_per_cell_heun_bitwise_against_oracle_3d():
    """3-D HLLD with TVD and Heun on the per-cell path: bit for bit the
    oracle's (no transcendental in this configuration)."""
    case("dropin per-cell 3-D 12x10x8 HLLD TVD Heun vs oracle")
    g = mesh.grid_3d(12, 10, 8)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.35, speed=3.0))
    ocfg = oracle.cartesian_config(recon=1, time_scheme=1)
    dt0 = oracle.Solver(fm, W, ocfg).dt_scan()
    s = oracle.Solver(fm, W, ocfg)
    s.dt_next = dt0
    assert s.run(8) == 0
    cp, t, dtn, pot = run_gpu(fm, W, gpu_cfg(ocfg), 8, dt0, solver.F_PER_CELL)
    np.testing.assert_array_equal(cp[:, :8], s.W[:, :8])
    np.testing.assert_array_equal(pot, s.potok)
    assert t == s.t and dtn == s.dt_next


def test_per_cell_falls_back_when_a_face_is_unlisted():
    """A boundary face missing from its cell's Cell_gran: the face pass still
    counts its CFL candidate, the per-cell form could not, so the flag falls
    back to two launches per stage and the bits stay those of that form."""
    case("dropin per-cell fallback, heliosphere 60x30 with an unlisted face")
    fm, W = helio_mesh(60, 30)
    f = int(np.flatnonzero(fm.gran_neighbour[:, 1] == -1)[5])
    c1 = fm.gran_neighbour[f, 0] - 1
    cg = fm.cell_gran.copy()
    cg[c1][cg[c1] == f + 1] = 0
    fm.cell_gran = cg
    ocfg = oracle.reference_config()
    a = run_gpu(fm, W, gpu_cfg(ocfg), 6, 2.0e-9)
    b = run_gpu(fm, W, gpu_cfg(ocfg), 6, 2.0e-9, solver.F_PER_CELL)
    assert_same_run(a, b)
