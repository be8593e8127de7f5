"""Argument errors at the C ABI (include/setmhd.h): empty and zero-size
inputs, calls before anything is uploaded, negative step counts and
non-positive time steps come back as SETMHD_ERR_ARG, host-side, without a
launch; a context that refused them steps afterwards bit for bit like a
fresh one.  Zero steps leave the state unchanged bit for bit, and a batch
whose every slot was removed steps as a no-op and takes new grids."""
import ctypes

import numpy as np
import pytest

from conftest import rel_err
from set_amd import mesh, solver

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fresh_run(g, W, nsteps):
    with solver.Context(solver.cartesian_config()) as c:
        c.upload_structured(g, W)
        c.dt_from_state()
        c.step(nsteps)
        return c.download()


def refused(call):
    with pytest.raises(solver.SetMHDError) as e:
        call()
    assert e.value.status == solver.ERR_ARG, e.value


def test_argument_errors_leave_the_context_usable():
    case("argument errors on a structured context, then 3 steps against a fresh context")
    g = mesh.grid_3d(20, 18, 16)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    box = np.array([g.x0, g.x1, g.y0, g.y1, g.z0, g.z1], np.float64)
    cp = np.ascontiguousarray(W, np.float64)
    L = solver.load()
    with solver.Context(solver.cartesian_config()) as c:
        # nothing uploaded yet
        refused(lambda: c.step(1))
        refused(lambda: c.set_dt(1e-3))
        refused(lambda: c.dt_from_state())
        refused(lambda: c.download())
        refused(lambda: c.batch_step(1))
        # zero-size grids and an empty face list
        for nx, ny, nz in ((0, 18, 16), (20, 0, 16), (20, 18, 0), (-1, 18, 16)):
            assert L.setmhd_upload_structured(c._h, nx, ny, nz, _pd(box), _pd(cp), 0.0) == solver.ERR_ARG
        assert L.setmhd_upload_mesh(c._h, 0, 0, *([None] * 10), None, 0.0) == solver.ERR_ARG
        # a valid upload after the refusals
        c.upload_structured(g, W)
        c.dt_from_state()
        refused(lambda: c.step(-1))
        refused(lambda: c.set_dt(0.0))
        refused(lambda: c.set_dt(-1e-3))
        refused(lambda: c.set_dt(float("nan")))
        before, t_before = c.download()
        c.step(0)
        after, t_after = c.download()
        assert np.array_equal(before, after) and t_before == t_after
        c.step(3)
        got, t_got = c.download()
    ref, t_ref = fresh_run(g, W, 3)
    assert t_got == t_ref
    assert np.array_equal(got, ref)


def test_batch_with_every_slot_removed_steps_as_a_no_op():
    case("batch: remove the only slot, step the empty batch, admit again, 2 steps against alone")
    g = mesh.Grid(40, 20)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.2))
    with solver.Context(solver.cartesian_config()) as c:
        s = c.batch_add(g, W)
        c.batch_step(2)
        c.batch_remove(s)
        c.batch_wait()  # the removed slot's memory is freed here
        c.batch_step(5)  # no grid left: nothing runs
        refused(lambda: c.batch_query(s))
        refused(lambda: c.batch_step(-1))
        s2 = c.batch_add(g, W)
        c.batch_step(2)
        q = c.batch_query(s2)
        got, t = c.batch_download(s2)
    assert q["steps"] == 2
    ref, t_ref = fresh_run(g, W, 2)
    assert t == pytest.approx(t_ref, rel=1e-14)
    e = rel_err(mesh.prim_to_cons(got[:, :8]), mesh.prim_to_cons(ref[:, :8]), "inf")
    assert np.all(e <= 1e-12), e
