"""Seeded synthetic workloads of the heliosphere set-up.

* batch_grids: 64 independent 2-D grids, sizes log-uniform from 128x64 to
  1024x512, per grid inflow speed U(1.5, 3.5), density U(0.05, 0.2), |B|
  U(0.1, 0.5) and tilt U(0, 90) degrees.
* serving_requests: a Poisson stream of solves (mean rate 20/s), 60 % 2-D
  (256x128 to 1024x512) and 40 % 3-D (64^3 to 192^3), the same parameter
  ranges.
* heliosphere_face_mesh: a half-plane heliosphere on SET's face-list arrays
  for the drop-in path at reference semantics.
"""
from dataclasses import dataclass

import numpy as np

from . import mesh


def _params(rng):
    return mesh.Inflow(rho=float(rng.uniform(0.05, 0.2)), u=-float(rng.uniform(1.5, 3.5)),
                       b=float(rng.uniform(0.1, 0.5)), tilt_deg=float(rng.uniform(0.0, 90.0)),
                       seed=int(rng.integers(1 << 30)))


def batch_grids(n=64, seed=20261203):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        s = float(np.exp(rng.uniform(0.0, np.log(8.0))))
        g = mesh.grid_2d(int(round(128 * s)), int(round(64 * s)))
        out.append((g, mesh.heliosphere_state(g, inflow=_params(rng))))
    return out


@dataclass
class Request:
    rid: int
    arrival: float        # seconds after the stream starts
    grid: mesh.Grid
    inflow: mesh.Inflow

    def state(self):
        return mesh.heliosphere_state(self.grid, inflow=self.inflow)

    @property
    def cells(self):
        return self.grid.n_cell


def serving_requests(n, rate=20.0, seed=7, frac_2d=0.6):
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        if rng.uniform() < frac_2d:
            s = float(np.exp(rng.uniform(0.0, np.log(4.0))))
            g = mesh.grid_2d(int(round(256 * s)), int(round(128 * s)))
        else:
            m = int(round(float(np.exp(rng.uniform(np.log(64.0), np.log(192.0))))))
            g = mesh.grid_3d(m, m, m)
        out.append(Request(i, t, g, _params(rng)))
    return out


def heliosphere_face_mesh(nx, ny, seed=5):
    """Half-plane heliosphere mesh as the reference uses it: y = 0 is the
    symmetry axis (Gran_type 5), the other sides are outer boundaries, the
    field is zero (w = B = 0 semantics) and the neutral-H moments
    Cell_par(9:16) are seeded positive values.  Returns (FaceMesh, Cell_par)."""
    g = mesh.grid_2d(nx, ny)
    fm = mesh.face_mesh(g)
    axis = (fm.gran_normal[:, 1] == -1.0) & (fm.gran_center[:, 1] == g.y0)
    fm.gran_type[axis] = mesh.TYPE_AXIS
    W = mesh.heliosphere_state(g, inflow=mesh.Inflow(b=0.0))
    rng = np.random.default_rng(seed)
    for base in (8, 12):
        W[:, base] = rng.uniform(0.05, 0.2, fm.n_cell)
        W[:, base + 1] = rng.uniform(-3.0, -2.0, fm.n_cell)
        W[:, base + 2] = rng.uniform(-0.2, 0.2, fm.n_cell)
        W[:, base + 3] = rng.uniform(0.5, 1.5, fm.n_cell)
    return fm, W
