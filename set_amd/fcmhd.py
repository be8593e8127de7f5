"""SET's FCMHD binary files: the solver's input (mesh + state) and its
restart/output file.

Input ("FCMHD_N.bin"), written by SET_Gelios (Setka.cpp:1240-1565) and read
by Set_Storage (Storage_module.f90:43-160): a raw little-endian stream
without record markers (FORM='BINARY'), arrays column-major:

    time f64 | N_cell i32, N_gran i32 | Cell_par f64[16*Nc] |
    Cell_center f64[3*Nc] | Cell_Volume f64[Nc] | Cell_gran i32[6*Nc] |
    Gran_normal f64[3*Nf] | Gran_square f64[Nf] | Gran_center f64[3*Nf] |
    Gran_neighbour i32[2*Nf] | Gran_neighbour_TVD i32[2*Nf] | 121.0 |
    Gran_type i32[Nf] | 122.0 | Gran_POTOK f64[9*Nf] | 123.0

Output ("FCMHD_N.M_out.bin"), Save_Storage (Storage_module.f90:163-189) and
read back by Fill_data (:191-209): time f64 | Cell_par f64[16*Nc] | 321.0.

The reference only prints the sentinels; here a wrong sentinel, a short
file or a value Set_Storage rejects (:92-156) raises FCMHDError.  Time is
kept on restart (the reference resets it to 0 in Set_CUDA,
cuf_kernel.cuf:353).
"""
import numpy as np

from .mesh import FaceMesh

F64 = np.dtype("<f8")
I32 = np.dtype("<i4")


class FCMHDError(ValueError):
    pass


class _Reader:
    def __init__(self, buf):
        self.buf = buf
        self.off = 0

    def take(self, dtype, count):
        n = dtype.itemsize * count
        if self.off + n > len(self.buf):
            raise FCMHDError(f"file ends at byte {len(self.buf)}, needed {self.off + n}")
        a = np.frombuffer(self.buf, dtype, count, self.off).copy()
        self.off += n
        return a

    def sentinel(self, value):
        v = self.take(F64, 1)[0]
        if v != value:
            raise FCMHDError(f"sentinel {value} expected at byte {self.off - 8}, found {v}")


def validate(cell_par, gran_neighbour, gran_neighbour_tvd, n_cell):
    """Set_Storage's checks (Storage_module.f90:92-156)."""
    for col in (8, 11, 15, 12):  # Cell_par(9), (12), (16), (13)
        if np.any(cell_par[:, col] < 0.0):
            raise FCMHDError(f"Cell_par({col + 1}) < 0")
    if np.any(cell_par[:, 0] < float(np.float32(0.000000001))):
        raise FCMHDError("rho < 1e-9")
    if np.any(gran_neighbour[:, 0] > n_cell) or np.any(gran_neighbour[:, 0] < 1):
        raise FCMHDError("Gran_neighbour(1) out of 1..N_cell")
    if np.any(gran_neighbour_tvd[:, 0] > n_cell):
        raise FCMHDError("Gran_neighbour_TVD(1) > N_cell")


def read_input(path):
    """Set_Storage: returns (time, FaceMesh, Cell_par (Nc,16), Gran_POTOK (Nf,9))."""
    r = _Reader(open(path, "rb").read())
    time = r.take(F64, 1)[0]
    nc, nf = (int(v) for v in r.take(I32, 2))
    if nc <= 0 or nf <= 0:
        raise FCMHDError(f"bad sizes N_cell={nc} N_gran={nf}")
    cell_par = r.take(F64, 16 * nc).reshape(nc, 16)
    cell_center = r.take(F64, 3 * nc).reshape(nc, 3)
    cell_volume = r.take(F64, nc)
    cell_gran = r.take(I32, 6 * nc).reshape(nc, 6)
    gran_normal = r.take(F64, 3 * nf).reshape(nf, 3)
    gran_square = r.take(F64, nf)
    gran_center = r.take(F64, 3 * nf).reshape(nf, 3)
    gran_nb = r.take(I32, 2 * nf).reshape(nf, 2)
    gran_tvd = r.take(I32, 2 * nf).reshape(nf, 2)
    r.sentinel(121.0)
    gran_type = r.take(I32, nf)
    r.sentinel(122.0)
    potok = r.take(F64, 9 * nf).reshape(nf, 9)
    r.sentinel(123.0)
    validate(cell_par, gran_nb, gran_tvd, nc)
    fm = FaceMesh(nc, nf, cell_center, cell_volume, cell_gran, gran_normal, gran_square,
                  gran_center, gran_nb, gran_tvd, gran_type)
    return float(time), fm, cell_par, potok


def write_input(path, fm: FaceMesh, cell_par, time=0.0, gran_potok=None):
    """Writes an input file in the producer's layout (Setka.cpp:1240-1565)."""
    nc, nf = fm.n_cell, fm.n_gran
    pot = np.zeros((nf, 9)) if gran_potok is None else np.asarray(gran_potok)
    with open(path, "wb") as f:
        f.write(np.array([time], F64).tobytes())
        f.write(np.array([nc, nf], I32).tobytes())
        for a, dt in ((cell_par, F64), (fm.cell_center, F64), (fm.cell_volume, F64),
                      (fm.cell_gran, I32), (fm.gran_normal, F64), (fm.gran_square, F64),
                      (fm.gran_center, F64), (fm.gran_neighbour, I32),
                      (fm.gran_neighbour_tvd, I32)):
            f.write(np.ascontiguousarray(a, dt).tobytes())
        f.write(np.array([121.0], F64).tobytes())
        f.write(np.ascontiguousarray(fm.gran_type, I32).tobytes())
        f.write(np.array([122.0], F64).tobytes())
        f.write(np.ascontiguousarray(pot, F64).tobytes())
        f.write(np.array([123.0], F64).tobytes())


def save_storage(path, time, cell_par):
    """Save_Storage (Storage_module.f90:163-189)."""
    with open(path, "wb") as f:
        f.write(np.array([time], F64).tobytes())
        f.write(np.ascontiguousarray(cell_par, F64).tobytes())
        f.write(np.array([321.0], F64).tobytes())


def fill_data(path, n_cell):
    """Fill_data (Storage_module.f90:191-209): (time, Cell_par)."""
    r = _Reader(open(path, "rb").read())
    time = r.take(F64, 1)[0]
    cell_par = r.take(F64, 16 * n_cell).reshape(n_cell, 16)
    r.sentinel(321.0)
    if r.off != len(r.buf):
        raise FCMHDError(f"{len(r.buf) - r.off} trailing bytes after the 321.0 sentinel")
    return float(time), cell_par
