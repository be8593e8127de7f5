"""Renumbering of face-list meshes (SETMHD_F_REORDER): the numbering the
library gives a mesh (setmhd_mesh_order, host only) and the renumbering
helper the GPU tests use, pinned on the CPU.

  * the order is the Z-order curve: on a 4x4 lattice it visits the cells in
    the textbook Morton sequence, and on scattered 3-D points it sorts by a
    bit-by-bit interleave written out here independently;
  * faces are grouped by the axis their normal is closest to, then follow
    the new place of their first cell;
  * the oracle on a randomly renumbered mesh gives the same bits, mapped
    back, as on the original mesh: the step does not depend on the
    numbering (reference semantics with TVD, Heun and sources, and 3-D
    HLLD), which is what lets the device hold its own order."""
import numpy as np
import pytest

import oracle
from set_amd import mesh, solver, workloads


def test_morton_order_on_a_lattice():
    g = mesh.Grid(4, 4, 1, 0.0, 4.0, 0.0, 4.0)
    fm = mesh.face_mesh(g)
    cp, gp = solver.mesh_order(fm)
    # (i, j) -> i + 4 j visited in Z order: x is the lowest interleaved bit
    assert cp.tolist() == [0, 1, 4, 5, 2, 3, 6, 7, 8, 9, 12, 13, 10, 11, 14, 15]
    assert sorted(gp.tolist()) == list(range(fm.n_gran))
    # the 20 x-faces (normals +-x) come first, then the 20 y-faces
    assert np.all(np.abs(fm.gran_normal[gp[:20], 0]) == 1.0)
    assert np.all(np.abs(fm.gran_normal[gp[20:], 1]) == 1.0)


def _interleave_key(q):
    """Morton key of 21-bit integer coordinates q (n x 3), bit by bit."""
    key = np.zeros(len(q), np.uint64)
    for b in range(21):
        for d in range(3):
            bit = (q[:, d].astype(np.uint64) >> np.uint64(b)) & np.uint64(1)
            key |= bit << np.uint64(3 * b + d)
    return key


def test_order_matches_bitwise_interleave():
    rng = np.random.default_rng(3)
    n = 500
    fm = mesh.face_mesh(mesh.grid_3d(5, 5, 5))
    centre = rng.uniform(-3.0, 7.0, (n, 3))
    centre[:7] = centre[7]  # ties keep the caller's order
    nb = np.stack([rng.integers(1, n + 1, 900), np.full(900, -1)], 1).astype(np.int32)
    normal = rng.normal(size=(900, 3))
    normal[:30] = [0.6, -0.6, 0.1]  # a tie between x and y goes to x
    normal /= np.linalg.norm(normal, axis=1)[:, None]
    fake = mesh.FaceMesh(n, 900, centre, np.ones(n), np.zeros((n, 6), np.int32), normal,
                         np.ones(900), np.zeros((900, 3)), nb, nb, np.ones(900, np.int32))
    cp, gp = solver.mesh_order(fake)
    lo, hi = centre.min(0), centre.max(0)
    q = np.clip((centre - lo) / (hi - lo), 0.0, 1.0) * float((1 << 21) - 1)
    key = _interleave_key(q.astype(np.int64))
    want = np.lexsort((np.arange(n), key))
    assert cp.tolist() == want.tolist()
    cinv = np.empty(n, np.int64)
    cinv[cp] = np.arange(n)
    first = cinv[nb[:, 0] - 1]
    axis = np.argmax(np.abs(normal), axis=1)  # first maximum: ties to the lower axis
    assert gp.tolist() == np.lexsort((np.arange(900), first, axis)).tolist()


def test_order_rejects_bad_neighbours():
    fm = mesh.face_mesh(mesh.grid_2d(3, 2))
    fm.gran_neighbour = fm.gran_neighbour.copy()
    fm.gran_neighbour[4, 0] = fm.n_cell + 1
    with pytest.raises(solver.SetMHDError):
        solver.mesh_order(fm)


def _shuffled(fm, seed):
    rng = np.random.default_rng(seed)
    return rng.permutation(fm.n_cell), rng.permutation(fm.n_gran)


@pytest.mark.parametrize("which", ["reference", "cartesian3d"])
def test_oracle_step_does_not_depend_on_the_numbering(which):
    if which == "reference":
        fm, W = workloads.heliosphere_face_mesh(40, 20)
        cfg, nsteps, dt0 = oracle.reference_config(), 6, 2.0e-9
    else:
        g = mesh.grid_3d(9, 8, 7)
        fm = mesh.face_mesh(g)
        W = mesh.heliosphere_state(g, wind=mesh.Wind(r_in=0.3))
        cfg, nsteps, dt0 = oracle.cartesian_config(), 5, None
    cp, gp = _shuffled(fm, 11)
    fr = mesh.renumber_face_mesh(fm, cp, gp)
    a = oracle.Solver(fm, W, cfg)
    b = oracle.Solver(fr, W[cp], cfg)
    a.dt_next = b.dt_next = dt0 if dt0 is not None else a.dt_scan()
    assert a.run(nsteps) == 0 and b.run(nsteps) == 0
    np.testing.assert_array_equal(b.W, a.W[cp])
    np.testing.assert_array_equal(b.potok, a.potok[gp])
    assert a.t == b.t and a.dt_next == b.dt_next


def test_renumbering_round_trip():
    fm, W = workloads.heliosphere_face_mesh(12, 6)
    cp, gp = _shuffled(fm, 2)
    fr = mesh.renumber_face_mesh(fm, cp, gp)
    cinv = np.argsort(cp)
    finv = np.argsort(gp)
    back = mesh.renumber_face_mesh(fr, cinv, finv)
    for k in ("cell_center", "cell_volume", "cell_gran", "gran_normal", "gran_square", "gran_center",
              "gran_neighbour", "gran_neighbour_tvd", "gran_type"):
        np.testing.assert_array_equal(getattr(back, k), getattr(fm, k))
