"""ctypes binding of libsetmhd.so (include/setmhd.h) and a host driver that
mirrors the reference's Fortran control flow.

Reference interface -> here:
  STORAGE arrays, Set_Storage / Save_Storage / Fill_data
      (Storage_module.f90:43-209)            -> set_amd.fcmhd
  Set_CUDA            (cuf_kernel.cuf:351-382) -> Context.upload_mesh / upload_structured
  time_step2 = 2e-9   (cuf_kernel.cuf:758)     -> Context.set_dt
  CUDA_START_MGD loop (cuf_kernel.cuf:774-925) -> Context.step / cuda_start_mgd
  Send_data_to_Host   (cuf_kernel.cuf:384-400) -> Context.download

There is no CPU fallback: loading fails loudly when the HIP library has not
been built, and every non-zero status raises SetMHDError.
"""
import ctypes
import os

import numpy as np

from . import mesh as _mesh

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsetmhd.so")

D = ctypes.c_double
I32 = ctypes.c_int32
PD = ctypes.POINTER(D)
PI = ctypes.POINTER(I32)

OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_RHO, ERR_REGION, ERR_FACE_RHO, ERR_UNSUPPORTED = range(8)
F_ZERO_BW, F_AXISYM, F_SOURCE_MK, F_HELIO_HACKS, F_FAST, F_REORDER, F_PER_CELL = 1, 2, 4, 8, 16, 32, 64
RIEMANN_LAX, RIEMANN_HLL, RIEMANN_HLLC, RIEMANN_HLLD = range(4)
TIME_EULER, TIME_HEUN = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

EXPORTS = (
    "setmhd_config_reference", "setmhd_config_cartesian", "setmhd_create", "setmhd_destroy",
    "setmhd_upload_mesh", "setmhd_upload_structured", "setmhd_set_dt", "setmhd_dt_from_state",
    "setmhd_step", "setmhd_step_timed", "setmhd_sync", "setmhd_download",
    "setmhd_download_potok", "setmhd_probe", "setmhd_download_async", "setmhd_download_wait",
    "setmhd_sizes", "setmhd_residual", "setmhd_last_error",
    "setmhd_batch_add", "setmhd_batch_remove", "setmhd_batch_step", "setmhd_batch_step_async",
    "setmhd_batch_poll", "setmhd_batch_wait", "setmhd_batch_query",
    "setmhd_batch_download", "setmhd_batch_autotune", "setmhd_batch_set_launch", "setmhd_mem_info",
    "setmhd_mesh_order", "setmhd_step_noalloc", "setmhd_batch_download_async",
    "setmhd_batch_download_collect", "setmhd_batch_download_view", "setmhd_ena_spectra",
)


class SetMHDError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"setmhd status {status}: {msg}")
        self.status = status


class Config(ctypes.Structure):
    """setmhd_config (include/setmhd.h)."""
    _fields_ = [("riemann", I32), ("recon", I32), ("time_scheme", I32), ("flags", I32),
                ("cfl", D), ("dt_cap", D), ("r_skip_face", D), ("r_hold_cell", D),
                ("precision", I32), ("device", I32), ("tile_x", I32), ("tile_y", I32),
                ("kchunk", I32)]


_lib = None


def load():
    """Load libsetmhd.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "(make -C set_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    L.setmhd_config_reference.argtypes = [ctypes.POINTER(Config)]
    L.setmhd_config_reference.restype = None
    L.setmhd_config_cartesian.argtypes = [ctypes.POINTER(Config), D, D]
    L.setmhd_config_cartesian.restype = None
    L.setmhd_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(Config)]
    L.setmhd_destroy.argtypes = [P]
    L.setmhd_destroy.restype = None
    L.setmhd_upload_mesh.argtypes = [P, I32, I32, PD, PD, PD, PI, PD, PD, PD, PI, PI, PI, PD, D]
    L.setmhd_upload_structured.argtypes = [P, I32, I32, I32, PD, PD, D]
    L.setmhd_set_dt.argtypes = [P, D]
    L.setmhd_dt_from_state.argtypes = [P, PD]
    L.setmhd_step.argtypes = [P, I32]
    L.setmhd_step_timed.argtypes = [P, I32, ctypes.POINTER(ctypes.c_float)]
    L.setmhd_sync.argtypes = [P, PD, PD]
    L.setmhd_download.argtypes = [P, PD, PD]
    L.setmhd_download_potok.argtypes = [P, PD]
    L.setmhd_probe.argtypes = [P, I32, PI, PD, PD]
    L.setmhd_download_async.argtypes = [P]
    L.setmhd_download_wait.argtypes = [P, PD, PD]
    L.setmhd_sizes.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.setmhd_residual.argtypes = [P, PD]
    L.setmhd_last_error.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t]
    I64P = ctypes.POINTER(ctypes.c_int64)
    L.setmhd_batch_add.argtypes = [P, I32, I32, I32, PD, PD, D, ctypes.POINTER(I32)]
    L.setmhd_batch_remove.argtypes = [P, I32]
    L.setmhd_batch_step.argtypes = [P, I32]
    L.setmhd_batch_set_launch.argtypes = [P, I32, I32, I32, I32]
    L.setmhd_batch_step_async.argtypes = [P, I32]
    L.setmhd_batch_wait.argtypes = [P]
    L.setmhd_batch_poll.argtypes = [P, ctypes.POINTER(I32)]
    L.setmhd_batch_query.argtypes = [P, I32, I64P, PD, PD, PD]
    L.setmhd_batch_download.argtypes = [P, I32, PD, PD]
    L.setmhd_batch_download_async.argtypes = [P, I32, ctypes.POINTER(I32)]
    L.setmhd_batch_download_collect.argtypes = [P, I32, I32, PD, PD, ctypes.POINTER(I32)]
    L.setmhd_batch_download_view.argtypes = [P, I32, ctypes.POINTER(PD), PD, ctypes.POINTER(I32)]
    L.setmhd_batch_autotune.argtypes = [P, I32, ctypes.POINTER(I32), ctypes.POINTER(I32),
                                        ctypes.POINTER(I32)]
    L.setmhd_mem_info.argtypes = [P, I64P, I64P]
    L.setmhd_mesh_order.argtypes = [I32, I32, PD, PD, PI, PI, PI]
    L.setmhd_ena_spectra.argtypes = [I32, I32, PD, PD, I32, PI, PI, I32, PD, I32, PD, PI]
    L.setmhd_step_noalloc.argtypes = [ctypes.POINTER(Config), I32, I32, I32, PD, ctypes.c_void_p,
                                      ctypes.c_void_p, D, ctypes.c_void_p, PD, PD]
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is not None and name not in ("setmhd_config_reference",
                                                  "setmhd_config_cartesian", "setmhd_destroy"):
            f.restype = ctypes.c_int
    _lib = L
    return L


def reference_config():
    """The reference's production semantics (setmhd_config_reference)."""
    c = Config()
    load().setmhd_config_reference(ctypes.byref(c))
    return c


def cartesian_config(cfl=0.4, r_hold=0.1, precision=PREC_FP64, time_scheme=TIME_EULER,
                     tile=(0, 0), device=0, kchunk=0):
    """First-order HLLD ideal MHD on a box (setmhd_config_cartesian)."""
    c = Config()
    load().setmhd_config_cartesian(ctypes.byref(c), float(cfl), float(r_hold))
    c.precision = precision
    c.time_scheme = time_scheme
    c.tile_x, c.tile_y = tile
    c.device = device
    c.kchunk = kchunk
    return c


def mesh_order(fm):
    """The device numbering SETMHD_F_REORDER gives face-list mesh fm
    (setmhd_mesh_order, host only): (cell_perm, gran_perm), 0-based, entry
    i being the caller's cell or face placed i-th."""
    cc = np.ascontiguousarray(fm.cell_center, np.float64)
    gn = np.ascontiguousarray(fm.gran_normal, np.float64)
    nb = np.ascontiguousarray(fm.gran_neighbour, np.int32)
    cp = np.zeros(fm.n_cell, np.int32)
    gp = np.zeros(fm.n_gran, np.int32)
    st = load().setmhd_mesh_order(fm.n_cell, fm.n_gran, _pd(cc), _pd(gn), _pi(nb), _pi(cp), _pi(gp))
    if st:
        raise SetMHDError(st, "mesh_order: first face neighbour out of range or empty mesh")
    return cp, gp


def step_noalloc(config, grid, d_in, d_out, dt, stream=None):
    """setmhd_step_noalloc: one forward-Euler step of a structured box whose
    conserved state (8 fields of grid.n_cell doubles) the caller keeps in
    device memory at address d_in; U^{n+1} goes to d_out.  Returns
    (dt_next, residual)."""
    box = np.array([grid.x0, grid.x1, grid.y0, grid.y1, grid.z0, grid.z1], np.float64)
    dtn, res = D(), D()
    st = load().setmhd_step_noalloc(ctypes.byref(config), grid.nx, grid.ny, grid.nz, _pd(box),
                                    ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), float(dt),
                                    ctypes.c_void_p(stream), ctypes.byref(dtn), ctypes.byref(res))
    if st != OK:
        raise SetMHDError(st, "step_noalloc failed")
    return dtn.value, res.value


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI)


class Context:
    """One setmhd_ctx: device state, stream and step loop of one problem."""

    def __init__(self, config):
        self._L = load()
        self._h = ctypes.c_void_p()
        self.config = config
        st = self._L.setmhd_create(ctypes.byref(self._h), ctypes.byref(config))
        if st != OK:
            raise SetMHDError(st, "setmhd_create failed")
        self.n_cell = 0

    def _check(self, st):
        if st != OK:
            buf = ctypes.create_string_buffer(512)
            self._L.setmhd_last_error(self._h, buf, 512)
            raise SetMHDError(st, buf.value.decode(errors="replace"))

    def close(self):
        if self._h:
            self._L.setmhd_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- Set_CUDA ---------------------------------------------------------
    def upload_mesh(self, fm, cell_par, gran_potok=None, t0=0.0):
        cp = np.ascontiguousarray(cell_par, np.float64)
        assert cp.shape == (fm.n_cell, 16)
        keep = [cp, np.ascontiguousarray(fm.cell_center), np.ascontiguousarray(fm.cell_volume),
                np.ascontiguousarray(fm.cell_gran, np.int32), np.ascontiguousarray(fm.gran_normal),
                np.ascontiguousarray(fm.gran_square), np.ascontiguousarray(fm.gran_center),
                np.ascontiguousarray(fm.gran_neighbour, np.int32),
                np.ascontiguousarray(fm.gran_neighbour_tvd, np.int32),
                np.ascontiguousarray(fm.gran_type, np.int32)]
        pot = None if gran_potok is None else np.ascontiguousarray(gran_potok, np.float64)
        self._check(self._L.setmhd_upload_mesh(
            self._h, fm.n_cell, fm.n_gran, _pd(keep[0]), _pd(keep[1]), _pd(keep[2]),
            _pi(keep[3]), _pd(keep[4]), _pd(keep[5]), _pd(keep[6]), _pi(keep[7]), _pi(keep[8]),
            _pi(keep[9]), None if pot is None else _pd(pot), float(t0)))
        self.n_cell = fm.n_cell
        self.n_gran = fm.n_gran

    def upload_structured(self, grid, cell_par, t0=0.0):
        cp = np.ascontiguousarray(cell_par, np.float64)
        assert cp.shape == (grid.n_cell, 16)
        box = np.array([grid.x0, grid.x1, grid.y0, grid.y1, grid.z0, grid.z1], np.float64)
        self._check(self._L.setmhd_upload_structured(self._h, grid.nx, grid.ny, grid.nz,
                                                     _pd(box), _pd(cp), float(t0)))
        self.n_cell = grid.n_cell

    # -- step loop ---------------------------------------------------------
    def set_dt(self, dt):
        self._check(self._L.setmhd_set_dt(self._h, float(dt)))

    def dt_from_state(self):
        d = D()
        self._check(self._L.setmhd_dt_from_state(self._h, ctypes.byref(d)))
        return d.value

    def step(self, nsteps=1):
        self._check(self._L.setmhd_step(self._h, int(nsteps)))

    def step_timed(self, nsteps):
        ms = ctypes.c_float()
        self._check(self._L.setmhd_step_timed(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def step_residual(self):
        r = D()
        self._check(self._L.setmhd_residual(self._h, ctypes.byref(r)))
        return r.value

    def sync(self):
        t, dt = D(), D()
        self._check(self._L.setmhd_sync(self._h, ctypes.byref(t), ctypes.byref(dt)))
        return t.value, dt.value

    # -- Send_data_to_Host -------------------------------------------------
    def download(self):
        cp = np.zeros((self.n_cell, 16))
        t = D()
        self._check(self._L.setmhd_download(self._h, _pd(cp), ctypes.byref(t)))
        return cp, t.value

    def download_potok(self):
        p = np.zeros((self.n_gran, 9))
        self._check(self._L.setmhd_download_potok(self._h, _pd(p)))
        return p

    def download_async(self):
        """Queues a snapshot of Cell_par and the time behind the queued
        steps; later steps may be queued before download_wait collects it."""
        self._check(self._L.setmhd_download_async(self._h))

    def download_wait(self):
        """The snapshot queued by download_async: (Cell_par array, t)."""
        cp = np.zeros((self.n_cell, 16))
        t = ctypes.c_double()
        self._check(self._L.setmhd_download_wait(self._h, _pd(cp), ctypes.byref(t)))
        return cp, t.value

    def probe(self, cells):
        """Cell_par(1:8) of the given 1-based cell numbers from the current
        state (one gather and one small copy); returns (n x 8 array, t)."""
        ids = np.ascontiguousarray(cells, np.int32).reshape(-1)
        out = np.zeros((ids.size, 8))
        t = ctypes.c_double()
        self._check(self._L.setmhd_probe(self._h, int(ids.size), ids.ctypes.data_as(PI), _pd(out),
                                         ctypes.byref(t)))
        return out, t.value

    # -- batches of structured grids ---------------------------------------
    def batch_add(self, grid, cell_par, t0=0.0):
        """Admits a structured grid; returns its slot id."""
        cp = np.ascontiguousarray(cell_par, np.float64)
        assert cp.shape == (grid.n_cell, 16)
        box = np.array([grid.x0, grid.x1, grid.y0, grid.y1, grid.z0, grid.z1], np.float64)
        slot = I32()
        self._check(self._L.setmhd_batch_add(self._h, grid.nx, grid.ny, grid.nz, _pd(box),
                                             _pd(cp), float(t0), ctypes.byref(slot)))
        self._slot_cells = getattr(self, "_slot_cells", {})
        self._slot_cells[slot.value] = grid.n_cell
        return slot.value

    def batch_remove(self, slot):
        """Retires a slot; its state stays readable until the next wait."""
        self._check(self._L.setmhd_batch_remove(self._h, int(slot)))

    def batch_step(self, nsteps=1):
        self._check(self._L.setmhd_batch_step(self._h, int(nsteps)))

    def batch_step_async(self, nsteps=1):
        """Queues nsteps steps and returns; see setmhd_batch_step_async."""
        self._check(self._L.setmhd_batch_step_async(self._h, int(nsteps)))

    def batch_wait(self):
        self._check(self._L.setmhd_batch_wait(self._h))

    def batch_poll(self):
        """True when the queued steps have finished (never blocks)."""
        d = I32()
        self._check(self._L.setmhd_batch_poll(self._h, ctypes.byref(d)))
        return bool(d.value)

    def batch_query(self, slot):
        n, t, dt, r = ctypes.c_int64(), D(), D(), D()
        self._check(self._L.setmhd_batch_query(self._h, int(slot), ctypes.byref(n), ctypes.byref(t),
                                               ctypes.byref(dt), ctypes.byref(r)))
        return {"steps": n.value, "t": t.value, "dt": dt.value, "residual": r.value}

    def batch_download(self, slot):
        cp = np.zeros((self._slot_cells[slot], 16))
        t = D()
        self._check(self._L.setmhd_batch_download(self._h, int(slot), _pd(cp), ctypes.byref(t)))
        return cp, t.value

    def batch_download_async(self, slot):
        """Queues a copy of the slot's state to pinned host memory; returns a
        ticket for batch_download_collect."""
        k = I32()
        self._check(self._L.setmhd_batch_download_async(self._h, int(slot), ctypes.byref(k)))
        self._tickets = getattr(self, "_tickets", {})
        self._tickets[k.value] = self._slot_cells[slot]
        return k.value

    def batch_download_collect(self, ticket, wait=True):
        """(Cell_par, t) of a queued copy, or None while it runs (wait=False)."""
        cp = np.zeros((self._tickets[ticket], 16))
        t, done = D(), I32()
        self._check(self._L.setmhd_batch_download_collect(self._h, int(ticket), int(bool(wait)), _pd(cp),
                                                          ctypes.byref(t), ctypes.byref(done)))
        if not done.value:
            return None
        del self._tickets[ticket]
        return cp, t.value

    def batch_download_view(self, ticket):
        """(Cell_par view, t) of a landed copy, reading the library's pinned
        buffer in place (valid until batch_download_release), or None while
        the copy runs; never blocks."""
        ptr, t, done = PD(), D(), I32()
        self._check(self._L.setmhd_batch_download_view(self._h, int(ticket), ctypes.byref(ptr),
                                                       ctypes.byref(t), ctypes.byref(done)))
        if not done.value:
            return None
        n = self._tickets[ticket]
        return np.ctypeslib.as_array(ptr, shape=(n, 16)), t.value

    def batch_download_release(self, ticket):
        """Frees a landed copy's ticket (its buffer goes back to the pool)."""
        done = I32()
        self._check(self._L.setmhd_batch_download_collect(self._h, int(ticket), 1, None, None,
                                                          ctypes.byref(done)))
        del self._tickets[ticket]

    def batch_set_launch(self, tile=(32, 8), blocks_per_cu=0, dim=0):
        """Fixes the batch launch of 2-D (dim 2), 3-D (dim 3) or all (0)
        grids; blocks_per_cu 0: one block per tile.  The tile shape applies
        to 3-D grids (2-D grids march row segments)."""
        self._check(self._L.setmhd_batch_set_launch(self._h, int(dim), int(tile[0]), int(tile[1]),
                                                    int(blocks_per_cu)))

    def batch_autotune(self, probe_steps=10):
        """Picks the fastest launch of the 2-D and of the 3-D grids; returns
        {2: (tile_x, tile_y, blocks_per_cu), 3: (...)} (blocks_per_cu 0: one
        block per tile)."""
        tx, ty, bpc = (I32 * 2)(), (I32 * 2)(), (I32 * 2)()
        self._check(self._L.setmhd_batch_autotune(self._h, int(probe_steps), tx, ty, bpc))
        return {2: (tx[0], ty[0], bpc[0]), 3: (tx[1], ty[1], bpc[1])}

    def mem_info(self):
        f, t = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._L.setmhd_mem_info(self._h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def cuda_start_mgd(self, all_step, dt_init=2.0e-9):
        """CUDA_START_MGD (cuf_kernel.cuf:402-949) without the text dumps:
        initial dt, all_step steps, download."""
        self.set_dt(dt_init)
        self.step(all_step)
        return self.download()


def conserved(cell_par):
    """Conserved variables (rho, rho u, rho v, rho w, E, bx, by, bz) of a
    Cell_par array, b in Heaviside-Lorentz units (comparison helper)."""
    return _mesh.prim_to_cons(np.asarray(cell_par)[:, :8])
