"""Probes return exactly what a download returns at the probed cells, on
both device paths and after odd and even step counts (the structured
Euler state alternates between two buffers)."""
import numpy as np
import pytest

from set_amd import mesh, solver, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,steps", [(0, 3), (0, 4), (1, 3)])
def test_probe_structured(precision, steps):
    print(f"[case] probe structured 3-D 40x36x20 precision={precision} steps={steps}", flush=True)
    g = mesh.grid_3d(40, 36, 20)
    W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
    cells = [1, 17, 40 * 36 + 5, g.n_cell]
    with solver.Context(solver.cartesian_config(precision=precision)) as c:
        c.upload_structured(g, W)
        c.dt_from_state()
        c.step(steps)
        pr, tp = c.probe(cells)
        cp, t = c.download()
    np.testing.assert_array_equal(pr, cp[np.array(cells) - 1, :8])
    assert tp == t


def test_probe_face_list_and_range():
    print("[case] probe drop-in heliosphere 60x30 reference semantics", flush=True)
    fm, W = workloads.heliosphere_face_mesh(60, 30)
    cells = np.array([1, 2, 900, fm.n_cell], np.int32)
    with solver.Context(solver.reference_config()) as c:
        c.upload_mesh(fm, W)
        c.set_dt(2.0e-9)
        c.step(5)
        pr, tp = c.probe(cells)
        cp, t = c.download()
        with pytest.raises(solver.SetMHDError):
            c.probe([fm.n_cell + 1])
    np.testing.assert_array_equal(pr, cp[cells - 1, :8])
    assert tp == t


@pytest.mark.parametrize("path", ["structured", "face-list"])
def test_download_async_snapshot(path):
    """A snapshot holds the state at the point it was queued, although more
    steps are queued before it is collected; one snapshot at a time."""
    print(f"[case] download_async snapshot, {path}", flush=True)
    if path == "structured":
        g = mesh.grid_3d(40, 36, 20)
        W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
        ctx = solver.Context(solver.cartesian_config())
    else:
        fm, W = workloads.heliosphere_face_mesh(60, 30)
        ctx = solver.Context(solver.reference_config())
    with ctx as c:
        if path == "structured":
            c.upload_structured(g, W)
            c.dt_from_state()
        else:
            c.upload_mesh(fm, W)
            c.set_dt(2.0e-9)
        c.step(3)
        ref, tref = c.download()
        c.download_async()
        with pytest.raises(solver.SetMHDError):
            c.download_async()
        c.step(4)  # queued behind the snapshot
        snap, tsnap = c.download_wait()
        after, tafter = c.download()
    np.testing.assert_array_equal(snap, ref)
    assert tsnap == tref and tafter > tref
    assert not np.array_equal(after[:, :8], ref[:, :8])
