"""Batches of structured grids (setmhd_batch_*): every grid in a batch
evolves exactly as it does alone, whatever else is admitted or retired
around it, and the autotuner leaves the state where it found it."""
import numpy as np
import pytest

from conftest import rel_err
from set_amd import mesh, solver

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def cons(cp):
    return mesh.prim_to_cons(cp[:, :8])


def alone(g, W, nsteps, precision=0):
    with solver.Context(solver.cartesian_config(precision=precision)) as c:
        c.upload_structured(g, W)
        c.dt_from_state()
        c.step(nsteps)
        cp, t = c.download()
    return cp, t


def grids():
    out = []
    for k, (nx, ny, nz) in enumerate([(96, 48, 1), (130, 70, 1), (200, 100, 1), (40, 36, 34), (33, 17, 1)]):
        g = mesh.Grid(nx, ny) if nz == 1 else mesh.grid_3d(nx, ny, nz)
        W = mesh.heliosphere_state(g, inflow=mesh.Inflow(seed=100 + k, u=-1.5 - 0.4 * k,
                                                        b=0.1 + 0.08 * k, tilt_deg=15 * k),
                                   wind=mesh.Wind(r_in=0.2 if nz == 1 else 0.3))
        out.append((g, W))
    return out


@pytest.mark.parametrize("precision,launch", [(0, ((32, 8), 2, 0)), (1, ((32, 8), 2, 0)),
                                              (0, ((32, 8), 0, 0)), (1, ((64, 4), 0, 0)),
                                              (0, ((16, 16), 64, 0)), (0, None), (1, None)])
def test_batch_equals_alone(precision, launch):
    case(f"batch of 5 (2-D + 3-D) vs alone, precision={precision}, launch={launch}")
    gs = grids()
    with solver.Context(solver.cartesian_config(precision=precision)) as c:
        if launch:
            c.batch_set_launch(*launch)  # else the defaults (one block per tile)
        slots = [c.batch_add(g, W) for g, W in gs]
        c.batch_step(37)
        res = [c.batch_download(s) for s in slots]
        q = [c.batch_query(s) for s in slots]
    for (g, W), (cp, t), qq in zip(gs, res, q):
        ref, tref = alone(g, W, 37, precision)
        e = rel_err(cons(cp), cons(ref), "inf")
        print(g.nx, g.ny, g.nz, "max rel Linf", e.max(), "t", t, tref)
        assert qq["steps"] == 37
        assert t == pytest.approx(tref, rel=1e-14)
        assert np.all(e <= 1e-12), e


def test_admission_and_retirement_mid_run():
    case("batch admit / retire mid-run")
    gs = grids()
    (gA, WA), (gB, WB), (gC, WC) = gs[0], gs[3], gs[1]
    with solver.Context(solver.cartesian_config()) as c:
        a = c.batch_add(gA, WA)
        c.batch_step(10)
        b = c.batch_add(gB, WB)
        c.batch_step(10)
        c.batch_remove(a)
        cc = c.batch_add(gC, WC)
        c.batch_step(5)
        cpB, _ = c.batch_download(b)
        cpC, _ = c.batch_download(cc)
        assert c.batch_query(b)["steps"] == 15 and c.batch_query(cc)["steps"] == 5
    for cp, (g, W), n in ((cpB, (gB, WB), 15), (cpC, (gC, WC), 5)):
        ref, _ = alone(g, W, n)
        e = rel_err(cons(cp), cons(ref), "inf")
        assert np.all(e <= 1e-12), e


def test_async_step_overlaps_admission_and_download():
    case("batch step_async with an admission and a retired download in flight")
    gs = grids()
    (gA, WA), (gB, WB), (gC, WC) = gs[0], gs[3], gs[2]
    with solver.Context(solver.cartesian_config()) as c:
        a = c.batch_add(gA, WA)
        b = c.batch_add(gB, WB)
        c.batch_step(10)
        c.batch_remove(a)            # readable until the next wait
        c.batch_step_async(7)        # b runs
        cc = c.batch_add(gC, WC)     # uploaded beside it, joins the next step
        cpA, _ = c.batch_download(a)
        qa = c.batch_query(a)
        qc = c.batch_query(cc)
        c.batch_wait()
        with pytest.raises(solver.SetMHDError):
            c.batch_download(a)      # freed by the wait
        c.batch_step(5)
        cpB, _ = c.batch_download(b)
        cpC, _ = c.batch_download(cc)
        assert qa["steps"] == 10 and qc["steps"] == 0
        assert c.batch_query(b)["steps"] == 22 and c.batch_query(cc)["steps"] == 5
    for cp, (g, W), n in ((cpA, (gA, WA), 10), (cpB, (gB, WB), 22), (cpC, (gC, WC), 5)):
        ref, _ = alone(g, W, n)
        e = rel_err(cons(cp), cons(ref), "inf")
        assert np.all(e <= 1e-12), (n, e)


def test_autotune_restores_state():
    case("batch autotune restores state")
    gs = grids()
    with solver.Context(solver.cartesian_config()) as c:
        slots = [c.batch_add(g, W) for g, W in gs]
        c.batch_step(4)
        tuned = c.batch_autotune(probe_steps=3)
        for tx, ty, bpc in tuned.values():
            assert (tx, ty) in {(32, 8), (16, 16), (64, 4), (16, 8)} and bpc in (0, 1, 2, 4, 8, 64)
        assert tuned[2][:2] == (32, 8)  # 2-D grids march rows whatever the tile
        assert all(c.batch_query(s)["steps"] == 4 for s in slots)
        c.batch_step(6)
        res = [c.batch_download(s)[0] for s in slots]
    for (g, W), cp in zip(gs, res):
        ref, _ = alone(g, W, 10)
        e = rel_err(cons(cp), cons(ref), "inf")
        assert np.all(e <= 1e-12), e


def test_batch_residual_matches_single():
    case("batch residual vs single-grid residual")
    g, W = grids()[2]
    with solver.Context(solver.cartesian_config()) as c:
        s = c.batch_add(g, W)
        c.batch_step(9)
        r_batch = c.batch_query(s)["residual"]
    with solver.Context(solver.cartesian_config()) as c:
        c.upload_structured(g, W)
        c.dt_from_state()
        c.step(8)
        r_single = c.step_residual()
    assert r_batch > 0 and r_batch == pytest.approx(r_single, rel=1e-9)


def test_query_and_download_of_a_slot_in_the_running_step():
    """A grid admitted before an asynchronous step is part of that step: a
    query or download issued at once waits for it (state and step count of
    the finished step, never a torn one)."""
    case("batch: query / download of a fresh slot while its first step runs")
    (gA, WA), (gB, WB) = grids()[1], grids()[3]
    with solver.Context(solver.cartesian_config()) as c:
        a = c.batch_add(gA, WA)
        b = c.batch_add(gB, WB)
        c.batch_step_async(7)
        cpA, tA = c.batch_download(a)
        qb = c.batch_query(b)
        qa = c.batch_query(a)
        cpB, tB = c.batch_download(b)
        c.batch_wait()
    assert qa["steps"] == 7 and qb["steps"] == 7
    for cp, t, (g, W) in ((cpA, tA, (gA, WA)), (cpB, tB, (gB, WB))):
        ref, tref = alone(g, W, 7)
        e = rel_err(cons(cp), cons(ref), "inf")
        assert np.all(e <= 1e-12), e
        assert t == pytest.approx(tref, rel=1e-14)


def test_step_error_survives_an_overlapping_admission():
    """A step that drives a density to zero (a CFL number of 1e4) raises the
    step's error word; an admission whose upload and first-dt scan run
    beside that step checks its own word, so batch_wait still reports the
    step's error."""
    case("batch: rho error of a running step, admission beside it, error reported by wait")
    (gA, WA), (gB, WB) = grids()[0], grids()[2]
    with solver.Context(solver.cartesian_config(cfl=1.0e4)) as c:
        c.batch_add(gA, WA)
        c.batch_step_async(1)
        c.batch_add(gB, WB)  # its scan runs with dt = 0: no error of its own
        with pytest.raises(solver.SetMHDError) as e:
            c.batch_wait()
        assert e.value.status == solver.ERR_RHO


def test_dt_from_state_after_a_large_dt():
    """dt_from_state's scan discards its update and runs it with dt = 0, so
    an earlier set_dt far beyond the CFL limit cannot make it report a
    density error for a sound state."""
    case("dt_from_state after set_dt(1e3)")
    g, W = grids()[2]
    with solver.Context(solver.cartesian_config()) as c:
        c.upload_structured(g, W)
        c.set_dt(1.0e3)
        dt = c.dt_from_state()
        assert 0.0 < dt < 1.0
        c.step(3)
        cp, _ = c.download()
    ref, _ = alone(g, W, 3)
    np.testing.assert_array_equal(cp, ref)


def test_async_download_view_and_collect():
    """setmhd_batch_download_async of retired slots while the next step runs:
    the copy read in place (download_view) and the copy collected into the
    caller's array both equal the synchronous download."""
    case("batch download_async: view and collect beside a running step")
    gs = grids()
    with solver.Context(solver.cartesian_config()) as c:
        slots = [c.batch_add(g, W) for g, W in gs[:3]]
        c.batch_step(6)
        ref = [c.batch_download(s) for s in slots[:2]]
        c.batch_remove(slots[0])
        c.batch_remove(slots[1])
        c.batch_step_async(5)  # slot 2 runs; the two retired copies go beside it
        k0 = c.batch_download_async(slots[0])
        k1 = c.batch_download_async(slots[1])
        got1 = c.batch_download_collect(k1, wait=True)
        r = None
        while r is None:
            r = c.batch_download_view(k0)
        view0 = np.array(r[0])
        t0 = r[1]
        c.batch_download_release(k0)
        c.batch_wait()
    np.testing.assert_array_equal(view0, ref[0][0])
    assert t0 == ref[0][1]
    np.testing.assert_array_equal(got1[0], ref[1][0])
    assert got1[1] == ref[1][1]
