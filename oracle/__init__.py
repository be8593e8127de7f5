"""ctypes view of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker / CPU baseline; the product
(set_amd/libsetmhd.so) never loads it.  See oracle.h for what each function
restates from the reference.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_chlld.so")

D = ctypes.c_double
I32 = ctypes.c_int32
PD = ctypes.POINTER(D)
PI = ctypes.POINTER(I32)

F_ZERO_BW, F_AXISYM, F_SOURCE_MK, F_HELIO_HACKS = 1, 2, 4, 8
F4 = lambda x: float(np.float32(x))  # default-REAL literal promoted to REAL(8)


class Config(ctypes.Structure):
    _fields_ = [("riemann", I32), ("recon", I32), ("time_scheme", I32), ("flags", I32),
                ("cfl", D), ("dt_cap", D), ("r_skip_face", D), ("r_hold_cell", D)]


class Mesh(ctypes.Structure):
    _fields_ = [("n_cell", I32), ("n_gran", I32),
                ("cell_center", PD), ("cell_volume", PD), ("cell_gran", PI),
                ("gran_normal", PD), ("gran_square", PD), ("gran_center", PD),
                ("gran_neighbour", PI), ("gran_neighbour_tvd", PI), ("gran_type", PI)]


class EnaMesh(ctypes.Structure):
    _fields_ = [("n_cell", I32), ("n_tri", I32), ("cell_center", PD), ("cell_par", PD), ("tri", PI)]


def reference_config():
    """The reference's production semantics: HLLC + LAX wedge, minmod TVD,
    Heun RK2, axisymmetric + charge-exchange sources, w = B = 0."""
    return Config(riemann=2, recon=1, time_scheme=1,
                  flags=F_ZERO_BW | F_AXISYM | F_SOURCE_MK | F_HELIO_HACKS,
                  cfl=F4(0.2), dt_cap=1.0e5, r_skip_face=F4(0.02), r_hold_cell=F4(0.05))


def cartesian_config(riemann=3, recon=0, time_scheme=0, cfl=0.4, r_hold=0.1):
    """Ideal MHD on a Cartesian box: first-order HLLD, forward Euler, no
    sources, cells inside r_hold held.  dt = cfl*dist/max(|SL|,|SR|), the
    reference's formula (cuf_solvers.cuf:429) with its 0.2 replaced by cfl."""
    return Config(riemann=riemann, recon=recon, time_scheme=time_scheme, flags=0,
                  cfl=cfl, dt_cap=1.0e5, r_skip_face=-1.0, r_hold_cell=r_hold)


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        L.orc_chlld.argtypes = [ctypes.c_int, D, D, D, D, PD, PD, PD, PD, PD, PD, ctypes.c_int]
        L.orc_chlld.restype = ctypes.c_int
        L.orc_minmod.argtypes = [D, D]
        L.orc_minmod.restype = D
        L.orc_linear.argtypes = [D] * 7
        L.orc_linear.restype = D
        L.orc_source_mk.argtypes = [PD, PD]
        L.orc_face_pass.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(Mesh), ctypes.c_int,
                                    PD, PD, PD, PD, ctypes.c_int]
        L.orc_cell_pass.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(Mesh), ctypes.c_int,
                                    D, PD, PD, PD, ctypes.c_int]
        L.orc_face_states.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(Mesh), ctypes.c_int,
                                      ctypes.c_int, PD, PD, PD, PD]
        L.orc_face_states.restype = ctypes.c_int
        L.orc_dt_scan.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(Mesh), PD, ctypes.c_int]
        L.orc_dt_scan.restype = D
        L.orc_run.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(Mesh), PD, PD, PD, PD, PD,
                              ctypes.c_int, ctypes.c_int]
        L.orc_ena_velocity1.argtypes = [D, D]
        L.orc_ena_velocity1.restype = D
        L.orc_ena_maxwell.argtypes = [D] * 8
        L.orc_ena_maxwell.restype = D
        L.orc_ena_f.argtypes = [ctypes.c_int, D]
        L.orc_ena_f.restype = D
        L.orc_ena_int1.argtypes = [D, D]
        L.orc_ena_int1.restype = D
        L.orc_ena_los.argtypes = [ctypes.POINTER(EnaMesh), D, ctypes.c_int, PD]
        L.orc_ena_los.restype = ctypes.c_int
        for f in ("orc_face_pass", "orc_cell_pass", "orc_run"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI)


def chlld(n_state, normal, q1, q2, w=0.0, n_disc=0):
    q1 = np.ascontiguousarray(q1, np.float64)
    q2 = np.ascontiguousarray(q2, np.float64)
    out = np.zeros(8)
    dl, dp, dc = D(), D(), D()
    st = lib().orc_chlld(int(n_state), float(normal[0]), float(normal[1]), float(normal[2]),
                         float(w), _pd(q1), _pd(q2), ctypes.byref(dl), ctypes.byref(dp),
                         ctypes.byref(dc), _pd(out), int(n_disc))
    return out, dl.value, dp.value, dc.value, st


def minmod(x, y):
    return lib().orc_minmod(float(x), float(y))


def linear(*a):
    return lib().orc_linear(*map(float, a))


def source_mk(cell_par16):
    cp = np.ascontiguousarray(cell_par16, np.float64)
    out = np.zeros(3)
    lib().orc_source_mk(_pd(cp), _pd(out))
    return out


def default_threads():
    """OMP_NUM_THREADS when set, else the host's cores, at most 16."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return n if n > 0 else max(1, min(16, os.cpu_count() or 1))


class Solver:
    """Holds a FaceMesh and the state arrays (Cell_par, Cell_par2,
    Gran_POTOK) and runs the oracle's step loop on them."""

    def __init__(self, fm, cell_par, config, nthreads=None):
        self.fm = fm
        self._keep = {k: np.ascontiguousarray(getattr(fm, k)) for k in
                      ("cell_center", "cell_volume", "cell_gran", "gran_normal", "gran_square",
                       "gran_center", "gran_neighbour", "gran_neighbour_tvd", "gran_type")}
        k = self._keep
        self.mesh = Mesh(fm.n_cell, fm.n_gran, _pd(k["cell_center"]), _pd(k["cell_volume"]),
                         _pi(k["cell_gran"]), _pd(k["gran_normal"]), _pd(k["gran_square"]),
                         _pd(k["gran_center"]), _pi(k["gran_neighbour"]),
                         _pi(k["gran_neighbour_tvd"]), _pi(k["gran_type"]))
        self.cfg = config
        self.W = np.ascontiguousarray(cell_par, np.float64).copy()
        assert self.W.shape == (fm.n_cell, 16)
        self.W2 = self.W.copy()
        self.potok = np.zeros((fm.n_gran, 9))
        self.t = 0.0
        self.dt_next = 2.0e-9  # cuf_kernel.cuf:758
        self.nthreads = nthreads or default_threads()

    def dt_scan(self):
        return lib().orc_dt_scan(ctypes.byref(self.cfg), ctypes.byref(self.mesh), _pd(self.W),
                                 self.nthreads)

    def run(self, nsteps):
        t, dtn = D(self.t), D(self.dt_next)
        st = lib().orc_run(ctypes.byref(self.cfg), ctypes.byref(self.mesh), _pd(self.W),
                           _pd(self.W2), _pd(self.potok), ctypes.byref(t), ctypes.byref(dtn),
                           int(nsteps), self.nthreads)
        self.t, self.dt_next = t.value, dtn.value
        return st

    def face_pass(self, stage, dt_min=1.0e15):
        d = D(dt_min)
        st = lib().orc_face_pass(ctypes.byref(self.cfg), ctypes.byref(self.mesh), int(stage),
                                 _pd(self.W), _pd(self.W2), _pd(self.potok), ctypes.byref(d),
                                 self.nthreads)
        return st, d.value

    def face_states(self, f, stage=0):
        """(status, q1, q2): the two primitive states face f (0-based) hands
        to chlld in the given stage."""
        q1, q2 = np.zeros(8), np.zeros(8)
        st = lib().orc_face_states(ctypes.byref(self.cfg), ctypes.byref(self.mesh), int(f), int(stage),
                                   _pd(self.W), _pd(self.W2), _pd(q1), _pd(q2))
        return st, q1, q2

    def cell_pass(self, stage, dt):
        return lib().orc_cell_pass(ctypes.byref(self.cfg), ctypes.byref(self.mesh), int(stage),
                                   float(dt), _pd(self.W), _pd(self.W2), _pd(self.potok),
                                   self.nthreads)


def ena_velocity1(u, cp):
    return lib().orc_ena_velocity1(float(u), float(cp))


def ena_maxwell(*a):
    return lib().orc_ena_maxwell(*map(float, a))


def ena_f(which, x):
    return lib().orc_ena_f(int(which), float(x))


def ena_int1(x, cp):
    return lib().orc_ena_int1(float(x), float(cp))


def ena_los(cell_center, cell_par, tri, the, n_v):
    """(jENA per velocity, samples inside the mesh) of one line of sight at
    angle `the` (radians from the x-axis) through the triangulated cell
    centres (SET_ENA.cpp:782-931)."""
    cc = np.ascontiguousarray(cell_center, np.float64)
    cp = np.ascontiguousarray(cell_par, np.float64)
    tr = np.ascontiguousarray(tri, np.int32)
    m = EnaMesh(len(cc), len(tr), _pd(cc), _pd(cp), _pi(tr))
    out = np.zeros(int(n_v))
    n = lib().orc_ena_los(ctypes.byref(m), float(the), int(n_v), _pd(out))
    return out, n
