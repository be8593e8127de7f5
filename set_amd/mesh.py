"""Host-side problem set-up: Cartesian grids in SET's face-list layout, and
the heliosphere initial states (supersonic radial wind inside r < r_in, a
uniform magnetised inflow elsewhere).

The face-list arrays are exactly the ones SET's solver reads from its
binary input (SET_Fortran_Cuda_MHD/Storage_module.f90:16-37, written by
SET_Gelios/source/Setka.cpp:1158-1567): column-major with the component
index fastest -- here C-ordered (N, k) numpy arrays -- 1-based cell and face
numbers, -1 for "no neighbour", 0 for "no face".  Face normals point from
neighbour 1 to neighbour 2; a boundary face has its one cell as neighbour 1
and -1 as neighbour 2, with an outward normal.  Gran_type 2 marks interior
faces and 4 the outer boundary (copy ghost), as the reference's producer
writes them (Setka.cpp:1525-1540).

A cell's face slots are ordered (x-, x+, y-, y+[, z-, z+]): the order in
which the fused structured kernel accumulates face fluxes, so both paths
sum in one order.
"""
from dataclasses import dataclass

import numpy as np

GAMMA = 5.0 / 3.0
CPI4 = 12.56637061435917295384          # Storage_module.f90:8
SPI4 = float(np.sqrt(CPI4))
NPAR = 16                               # Storage_module.f90:13

TYPE_INTERIOR = 2
TYPE_OUTER = 4
TYPE_AXIS = 5


@dataclass
class FaceMesh:
    """SET's mesh arrays (Storage_module.f90:16-37), 1-based numbering."""
    n_cell: int
    n_gran: int
    cell_center: np.ndarray         # (Nc, 3) f64
    cell_volume: np.ndarray         # (Nc,)   f64
    cell_gran: np.ndarray           # (Nc, 6) i32
    gran_normal: np.ndarray         # (Nf, 3) f64
    gran_square: np.ndarray         # (Nf,)   f64
    gran_center: np.ndarray         # (Nf, 3) f64
    gran_neighbour: np.ndarray      # (Nf, 2) i32
    gran_neighbour_tvd: np.ndarray  # (Nf, 2) i32
    gran_type: np.ndarray           # (Nf,)   i32


@dataclass
class Grid:
    """A uniform Cartesian box of nx*ny*nz cells (nz == 1 with dim == 2)."""
    nx: int
    ny: int
    nz: int = 1
    x0: float = -1.0
    x1: float = 1.0
    y0: float = 0.0
    y1: float = 1.0
    z0: float = 0.0
    z1: float = 1.0

    @property
    def dim(self):
        return 2 if self.nz == 1 else 3

    @property
    def n_cell(self):
        return self.nx * self.ny * self.nz

    @property
    def dx(self):
        return (self.x1 - self.x0) / self.nx

    @property
    def dy(self):
        return (self.y1 - self.y0) / self.ny

    @property
    def dz(self):
        return (self.z1 - self.z0) / self.nz if self.dim == 3 else 1.0

    def axes(self):
        i = np.arange(self.nx)
        j = np.arange(self.ny)
        k = np.arange(self.nz)
        xc = self.x0 + (i + 0.5) * self.dx
        yc = self.y0 + (j + 0.5) * self.dy
        zc = self.z0 + (k + 0.5) * self.dz if self.dim == 3 else np.zeros(1)
        return xc, yc, zc

    def cell_centers(self):
        """(Nc, 3) centres, x fastest (cell number c = i + nx*(j + ny*k))."""
        xc, yc, zc = self.axes()
        Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
        return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def grid_2d(nx, ny, box=(-1.0, 1.0, 0.0, 1.0)):
    return Grid(nx, ny, 1, box[0], box[1], box[2], box[3])


def grid_3d(nx, ny, nz, box=(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)):
    return Grid(nx, ny, nz, *box)


def face_mesh(g: Grid, boundary_type=TYPE_OUTER) -> FaceMesh:
    """Express grid g in SET's face-list layout (see module docstring)."""
    nx, ny, nz, dim = g.nx, g.ny, g.nz, g.dim
    nc = g.n_cell
    dx, dy, dz = g.dx, g.dy, g.dz
    xc, yc, zc = g.axes()
    xf = g.x0 + np.arange(nx + 1) * dx
    yf = g.y0 + np.arange(ny + 1) * dy
    zf = g.z0 + np.arange(nz + 1) * dz if dim == 3 else np.zeros(2)

    def cid(i, j, k):
        return 1 + i + nx * (j + ny * k)

    area_x = dy * dz if dim == 3 else dy
    area_y = dx * dz if dim == 3 else dx
    area_z = dx * dy
    blocks = []  # (centre, normal, square, nb, tvd, type) per axis

    # x-faces: f(i, j, k) = i + (nx+1)*(j + ny*k), i = 0..nx
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx + 1), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    cen = np.stack([xf[I], yc[J], zc[K] if dim == 3 else np.zeros_like(I, float)], 1)
    blocks.append(_axis_block(I, nx, lambda a: cid(a, J, K), 0, cen, area_x, boundary_type))
    # y-faces: f = i + nx*(j + (ny+1)*k), j = 0..ny
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny + 1), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    cen = np.stack([xc[I], yf[J], zc[K] if dim == 3 else np.zeros_like(I, float)], 1)
    blocks.append(_axis_block(J, ny, lambda a: cid(I, a, K), 1, cen, area_y, boundary_type))
    if dim == 3:
        K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny), np.arange(nx), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        cen = np.stack([xc[I], yc[J], zf[K]], 1)
        blocks.append(_axis_block(K, nz, lambda a: cid(I, J, a), 2, cen, area_z, boundary_type))

    gran_center = np.concatenate([b[0] for b in blocks])
    gran_normal = np.concatenate([b[1] for b in blocks])
    gran_square = np.concatenate([b[2] for b in blocks])
    gran_nb = np.concatenate([b[3] for b in blocks]).astype(np.int32)
    gran_tvd = np.concatenate([b[4] for b in blocks]).astype(np.int32)
    gran_type = np.concatenate([b[5] for b in blocks]).astype(np.int32)
    nfx = (nx + 1) * ny * nz
    nfy = nx * (ny + 1) * nz

    # cell -> face slots (1-based)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    fxm = 1 + I + (nx + 1) * (J + ny * K)
    fxp = fxm + 1
    fym = 1 + nfx + I + nx * (J + (ny + 1) * K)
    fyp = fym + nx
    cell_gran = np.zeros((nc, 6), np.int32)
    cell_gran[:, 0], cell_gran[:, 1], cell_gran[:, 2], cell_gran[:, 3] = fxm, fxp, fym, fyp
    if dim == 3:
        fzm = 1 + nfx + nfy + I + nx * (J + ny * K)
        cell_gran[:, 4], cell_gran[:, 5] = fzm, fzm + nx * ny
    vol = dx * dy * dz if dim == 3 else dx * dy
    return FaceMesh(
        n_cell=nc, n_gran=len(gran_type),
        cell_center=np.ascontiguousarray(g.cell_centers()),
        cell_volume=np.full(nc, vol),
        cell_gran=cell_gran,
        gran_normal=np.ascontiguousarray(gran_normal),
        gran_square=np.ascontiguousarray(gran_square),
        gran_center=np.ascontiguousarray(gran_center),
        gran_neighbour=np.ascontiguousarray(gran_nb),
        gran_neighbour_tvd=np.ascontiguousarray(gran_tvd),
        gran_type=gran_type)


def renumber_face_mesh(fm: FaceMesh, cell_perm, gran_perm) -> FaceMesh:
    """fm with its cells and faces renumbered: new cell i is old cell
    cell_perm[i], new face f is old face gran_perm[f] (0-based).  Cell and
    face numbers in the connectivity are remapped, the -1 / 0 sentinels
    kept.  A state goes along as cell_par[cell_perm], a face array as
    gran_potok[gran_perm]."""
    cp = np.asarray(cell_perm, np.int64)
    gp = np.asarray(gran_perm, np.int64)
    cinv = np.empty_like(cp)
    cinv[cp] = np.arange(cp.size)
    finv = np.empty_like(gp)
    finv[gp] = np.arange(gp.size)

    def remap(a, inv):
        return np.where(a >= 1, inv[np.maximum(a, 1) - 1] + 1, a).astype(np.int32)

    return FaceMesh(
        n_cell=fm.n_cell, n_gran=fm.n_gran,
        cell_center=np.ascontiguousarray(fm.cell_center[cp]),
        cell_volume=np.ascontiguousarray(fm.cell_volume[cp]),
        cell_gran=remap(fm.cell_gran[cp], finv),
        gran_normal=np.ascontiguousarray(fm.gran_normal[gp]),
        gran_square=np.ascontiguousarray(fm.gran_square[gp]),
        gran_center=np.ascontiguousarray(fm.gran_center[gp]),
        gran_neighbour=remap(fm.gran_neighbour[gp], cinv),
        gran_neighbour_tvd=remap(fm.gran_neighbour_tvd[gp], cinv),
        gran_type=np.ascontiguousarray(fm.gran_type[gp]).astype(np.int32))


def _axis_block(A, n, cell_at, axis, centre, area, boundary_type):
    """Faces along one axis.  A = face position 0..n along the axis."""
    m = len(A)
    normal = np.zeros((m, 3))
    normal[:, axis] = 1.0
    nb = np.full((m, 2), -1, np.int64)
    tvd = np.full((m, 2), -1, np.int64)
    ftype = np.full(m, TYPE_INTERIOR, np.int64)
    interior = (A > 0) & (A < n)
    lo = A == 0
    hi = A == n
    nb[interior, 0] = cell_at(A - 1)[interior]
    nb[interior, 1] = cell_at(np.minimum(A, n - 1))[interior]
    ok1 = interior & (A - 2 >= 0)
    ok2 = interior & (A + 1 <= n - 1)
    tvd[ok1, 0] = cell_at(np.maximum(A - 2, 0))[ok1]
    tvd[ok2, 1] = cell_at(np.minimum(A + 1, n - 1))[ok2]
    nb[lo, 0] = cell_at(np.zeros_like(A))[lo]
    normal[lo, axis] = -1.0
    nb[hi, 0] = cell_at(np.full_like(A, n - 1))[hi]
    ftype[lo | hi] = boundary_type
    return centre, normal, np.full(m, area), nb, tvd, ftype


# ----------------------------------------------------------------------------
# initial states

def counter_uniform(seed, index):
    """Counter-based U[0,1): splitmix64 of (seed, cell index).  The same draw
    for a cell whatever the order or device that evaluates it."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
             + np.asarray(index, np.uint64) * np.uint64(0xD1B54A32D192ED03)
             + np.uint64(0x632BE59BD9B4E019))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Inflow:
    rho: float = 0.1
    u: float = -2.5
    p: float = 0.1
    b: float = 0.3          # |B| in SET's stored (Gaussian) units, Cell_par(6:8)
    tilt_deg: float = 30.0  # angle of B from +x in the x-y plane
    noise: float = 0.01     # +-1 % uniform density noise
    seed: int = 1234


@dataclass
class Wind:
    r_in: float = 0.1       # cells with |centre| < r_in hold the wind
    speed: float = 10.0     # radial |u|
    p_over_rho: float = 0.01


def heliosphere_state(g: Grid, inflow: Inflow = Inflow(), wind: Wind = Wind()):
    """Cell_par (Nc, 16): rho,u,v,w,p,Bx,By,Bz then 8 neutral moments (0).

    Wind cells (|centre| < r_in) get rho = 1/r^2 and u = speed * centre/r
    with r floored at half the smallest cell size, so a cell centred on
    the origin (odd cell counts) has finite density and zero velocity."""
    c = g.cell_centers()
    nc = len(c)
    W = np.zeros((nc, NPAR))
    idx = np.arange(nc, dtype=np.uint64)
    noise = (2.0 * counter_uniform(inflow.seed, idx) - 1.0) * inflow.noise
    th = np.deg2rad(inflow.tilt_deg)
    W[:, 0] = inflow.rho * (1.0 + noise)
    W[:, 1] = inflow.u
    W[:, 4] = inflow.p
    W[:, 5] = inflow.b * np.cos(th)
    W[:, 6] = inflow.b * np.sin(th)
    r = np.sqrt(c[:, 0] ** 2 + c[:, 1] ** 2 + c[:, 2] ** 2)
    inner = r < wind.r_in
    h = 0.5 * min(g.dx, g.dy, g.dz if g.dim == 3 else g.dx)
    rr = np.maximum(r[inner], h)
    W[inner, 0] = 1.0 / rr ** 2
    W[inner, 1:4] = wind.speed * c[inner] / rr[:, None]
    W[inner, 4] = wind.p_over_rho * W[inner, 0]
    W[inner, 5:8] = 0.0
    if g.dim == 2:
        W[:, 3] = 0.0
        W[:, 7] = 0.0
    return W


def prim_to_cons(W):
    """Cell_par primitives (N, >=8), stored Gaussian B -> conserved
    (rho, rho u, rho v, rho w, E, bx, by, bz) with b = B/sqrt(4 pi) and
    E = p/(gamma-1) + rho |u|^2 / 2 + |b|^2 / 2."""
    W = np.asarray(W)
    rho = W[:, 0]
    u = W[:, 1:4]
    b = W[:, 5:8] / SPI4
    U = np.empty((len(W), 8))
    U[:, 0] = rho
    U[:, 1:4] = rho[:, None] * u
    U[:, 4] = W[:, 4] / (GAMMA - 1.0) + 0.5 * rho * (u * u).sum(1) + 0.5 * (b * b).sum(1)
    U[:, 5:8] = b
    return U


def cons_to_prim(U):
    """Inverse of prim_to_cons; returns (N, 8) primitives with Gaussian B."""
    U = np.asarray(U)
    rho = U[:, 0]
    u = U[:, 1:4] / rho[:, None]
    b = U[:, 5:8]
    W = np.empty((len(U), 8))
    W[:, 0] = rho
    W[:, 1:4] = u
    W[:, 4] = (GAMMA - 1.0) * (U[:, 4] - 0.5 * rho * (u * u).sum(1) - 0.5 * (b * b).sum(1))
    W[:, 5:8] = b * SPI4
    return W
