"""Text outputs of SET's Fortran solver, in its own formats and units.

* write_state: x, y, rho, u, v, p, |V| per cell (Print, FCMHD.f90:6-50),
  Tecplot text with the format '(6(ES15.7, 2X))', whose format reversion puts
  |V| on a line of its own.
* write_density: x, y, rho per cell with a ZONE line carrying the solution
  time (Print_mult, FCMHD.f90:52-96), one file per dump.
* append_probes: one line 'T rho_1 ... rho_n' of a probe series
  (write_T_rho, FCMHD.f90:98-119, format '(18es23.15)').

Values are scaled to SET's dimensional units as the reference scales them:
lengths x 319.319, density x 0.073, velocity x 10.3565, pressure x 1.30962,
time x 146.262 (cuf_kernel.cuf:887,915).  The reference writes these as
default-real (single-precision) literals in double-precision products, so
the factors here are their single-precision values.  The state comes from
Context.download (Cell_par, 16 per cell) or Context.probe.
"""
import numpy as np

SCALE_X, SCALE_RHO, SCALE_U, SCALE_P, SCALE_T = (float(np.float32(v)) for v in
                                                 (319.319, 0.073, 10.3565, 1.30962, 146.262))


def _es(v, w, d):
    """Fortran ESw.d of one value (two-digit exponents)."""
    return f"{v:{w}.{d}E}"


def _records(values, per_record, w, d, gap):
    """Fortran list of values under '(per_record(ESw.d, gapX))' with format
    reversion: per_record values per line, each followed by gap blanks
    (trailing blanks dropped, as a record ends at its last character)."""
    lines = []
    for i in range(0, len(values), per_record):
        lines.append("".join(_es(v, w, d) + " " * gap for v in values[i:i + per_record]).rstrip())
    return lines


def write_state(path, cell_center, cell_par):
    """Print: TITLE/VARIABLES header, then x, y, rho, u, v, p and |V| per cell."""
    cc = np.asarray(cell_center, np.float64)
    cp = np.asarray(cell_par, np.float64)
    with open(path, "w") as f:
        f.write('TITLE = "HP"  VARIABLES = x, y, rho, u, v, p, |V|\n')
        for i in range(cp.shape[0]):
            x, y = cc[i, 0] * SCALE_X, cc[i, 1] * SCALE_X
            rho = cp[i, 0] * SCALE_RHO
            u, v = cp[i, 1] * SCALE_U, cp[i, 2] * SCALE_U
            p = cp[i, 4] * SCALE_P
            vals = [x, y, rho, u, v, p, float(np.sqrt(u * u + v * v))]
            f.write("\n".join(_records(vals, 6, 15, 7, 2)) + "\n")


def write_density(path, cell_center, cell_par, time):
    """Print_mult: header with the zone's solution time, then x, y, rho per cell.
    time is in code units (scaled by 146.262 here)."""
    cc = np.asarray(cell_center, np.float64)
    cp = np.asarray(cell_par, np.float64)
    t = time * SCALE_T
    with open(path, "w") as f:
        f.write('TITLE = "HP"\n')
        f.write("VARIABLES = x, y, rho\n")
        f.write('ZONE T="HP", STRANDID=1, SOLUTIONTIME=' + _es(t, 12, 5) + "\n")
        for i in range(cp.shape[0]):
            vals = [cc[i, 0] * SCALE_X, cc[i, 1] * SCALE_X, cp[i, 0] * SCALE_RHO]
            f.write(_records(vals, 3, 15, 7, 2)[0] + "\n")


def append_probes(path, time, rho):
    """write_T_rho: appends 'T rho_1 ... rho_n' (time scaled by 146.262,
    densities by 0.073), 18 values per line."""
    vals = [time * SCALE_T] + [float(r) * SCALE_RHO for r in np.asarray(rho, np.float64).reshape(-1)]
    with open(path, "a") as f:
        f.write("\n".join("".join(_es(v, 23, 15) for v in vals[i:i + 18]) for i in range(0, len(vals), 18))
                + "\n")
