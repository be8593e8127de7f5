"""ENA line-of-sight spectra of a heliosphere solution (SET's
post-processor, SET_ENA/SET_ENA.cpp:544-956) on the GPU.

    tri, nb = ena.triangulate(cell_center)
    j, n = ena.spectra(cell_center, cell_par, angles, n_v=1000, tri=(tri, nb))

The reference triangulates the cell centres with CGAL's Delaunay
triangulation (SET_ENA.cpp:242-257); here scipy's (Qhull) does it, the same
triangulation wherever no four centres are cocircular.  `spectra` runs
setmhd_ena_spectra (libsetmhd.so); the text writers follow the reference's
output files: the spectrum of one line of sight (Culc_ENA_the,
SET_ENA.cpp:934-952) and the intensities at fixed energies per angle
(Culc_ENA, :700-779), with the numbers at the stream's default precision
(6 significant digits, printf's %g) as the reference writes them.  `run`
and `run_map` are the reference's main (:958-975): the FCMHD input file for
the cell centres, the solver's output file for the state, then one line of
sight or the angle map.
"""
import ctypes

import numpy as np

from . import solver

AE = 0.00313166       # 1 AU in the solution's length unit (SET_ENA.cpp:791)
KMS = 0.0963358       # 1 km/s in its speed unit (:792)
FLUX = 6.75186e7      # output factor of the intensities (:700, :934)
# energies (eV) of the angle map (:729-771)
MAP_ENERGIES = (15.0, 29.0, 55.0, 110.0, 209.0, 439.0, 872.0, 1821.0,
                450.0, 710.0, 1110.0, 1740.0, 2730.0, 4290.0)


def triangulate(cell_center):
    """Delaunay triangles of the cell centres (x, y) and their neighbours:
    (tri, nb), tri[t] three 0-based cell numbers counterclockwise, nb[t, k]
    the triangle across the edge opposite vertex k (-1 on the hull)."""
    from scipy.spatial import Delaunay
    pts = np.ascontiguousarray(np.asarray(cell_center)[:, :2], np.float64)
    d = Delaunay(pts)
    tri = d.simplices.astype(np.int32)
    nb = d.neighbors.astype(np.int32)
    p = pts[tri]
    cw = ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
          - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1])) < 0
    tri[cw] = tri[cw][:, [0, 2, 1]]  # counterclockwise, as CGAL's faces
    nb[cw] = nb[cw][:, [0, 2, 1]]
    return np.ascontiguousarray(tri), np.ascontiguousarray(nb)


def map_angles():
    """The angles of Culc_ENA's map (:549): 0.008 rad, then repeatedly plus
    pi/180 while below pi - 0.008, accumulated as the reference's loop
    accumulates them -- 180 angles."""
    out, th = [], 0.008
    while th < np.pi - 0.008:
        out.append(th)
        th = th + np.pi / 180.0
    return np.array(out)


def speeds(n_v):
    """The particle speeds of a spectrum: v_i = vL + (i + 1/2)(vR - vL)/n_v,
    vL = 1 km/s, vR = 1000 km/s (:815-836)."""
    vl, vr = 1 * KMS, 1000 * KMS
    return vl + (np.arange(n_v) + 0.5) * ((vr - vl) / n_v)


def energies_ev(n_v):
    """velocity_to_energy of the speeds (:408-412), eV."""
    x = speeds(n_v) / KMS
    return x * x * 5.226 * 0.001


def spectra(cell_center, cell_par, angles, n_v=1000, tri=None, device=0):
    """(jENA [n_angle, n_v] before the output factor, samples per angle)."""
    if tri is None:
        tri = triangulate(cell_center)
    t, nb = tri
    cc = np.ascontiguousarray(cell_center, np.float64)
    cp = np.ascontiguousarray(cell_par, np.float64)
    ang = np.ascontiguousarray(np.atleast_1d(angles), np.float64)
    t = np.ascontiguousarray(t, np.int32)
    nb = np.ascontiguousarray(nb, np.int32)
    assert cc.shape == (len(cp), 3) and cp.shape[1] == 16
    out = np.zeros((len(ang), int(n_v)))
    smp = np.zeros(len(ang), np.int32)
    PD, PI = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
    st = solver.load().setmhd_ena_spectra(
        int(device), len(cc), cc.ctypes.data_as(PD), cp.ctypes.data_as(PD), len(t), t.ctypes.data_as(PI),
        nb.ctypes.data_as(PI), len(ang), ang.ctypes.data_as(PD), int(n_v), out.ctypes.data_as(PD),
        smp.ctypes.data_as(PI))
    if st != solver.OK:
        raise solver.SetMHDError(st, "ena_spectra failed")
    return out, smp


def linear_interpolate(x, xs, ys):
    """linear_interpolate (:373-406): clamped at both ends."""
    if x <= xs[0]:
        return ys[0]
    if x >= xs[-1]:
        return ys[-1]
    i = int(np.searchsorted(xs, x, side="right")) - 1
    return ys[i] + (ys[i + 1] - ys[i]) * (x - xs[i]) / (xs[i + 1] - xs[i])


def angle_map(jena, n_v, fix_high_energies=False):
    """Intensities (times the output factor) at MAP_ENERGIES for each angle,
    as Culc_ENA writes them (:727-776).  The reference evaluates its last
    six columns (450-4290 eV) at 1821 eV; fix_high_energies=True takes each
    column's own energy instead."""
    e = energies_ev(n_v)
    out = np.zeros((len(jena), len(MAP_ENERGIES)))
    for a, row in enumerate(jena):
        for k, en in enumerate(MAP_ENERGIES):
            at = en if (k < 8 or fix_high_energies) else MAP_ENERGIES[7]
            out[a, k] = linear_interpolate(at, e, row) * FLUX
    return out


def _num(x, exact):
    return repr(float(x)) if exact else "%g" % x


def write_spectrum(path, jena_row, exact=False):
    """The spectrum file of one line of sight (Culc_ENA_the, :934-952):
    energy (keV), speed (km/s), intensity per line.  exact=True writes every
    digit of each double instead of the reference's 6."""
    v = speeds(len(jena_row))
    with open(path, "w") as f:
        for vi, ji in zip(v, jena_row):
            e = (vi / KMS) * (vi / KMS) * 5.226 * 0.000001
            f.write(f"{_num(e, exact)} {_num(vi / KMS, exact)} {_num(ji * FLUX, exact)}\n")


def write_angle_map(path, angles, amap, exact=False):
    """Culc_ENA's jENA_angle_test.txt layout (:773-776): the angle, then the
    14 intensities."""
    with open(path, "w") as f:
        for th, row in zip(angles, amap):
            f.write(" ".join([_num(th, exact)] + [_num(x, exact) for x in row]) + "\n")


def read_solution(mesh_file, state_file):
    """Set_Storage then ReadUpdatedStorage (:115-240): the cell centres from
    the FCMHD input file and Cell_par from the solver's output file (time,
    16 x N_cell, the 321.0 sentinel).  Returns (cell_center, cell_par, time)."""
    from . import fcmhd
    _, fm, _, _ = fcmhd.read_input(mesh_file)
    t, cp = fcmhd.fill_data(state_file, fm.n_cell)
    return fm.cell_center, cp, t


def run(mesh_file, state_file, the, name, n_v=1000, out_dir=".", device=0):
    """main + Culc_ENA_the (:958-975, :782-956): the spectrum along the line
    of sight at angle `the` (rad), written to <out_dir>/jENA_<name>.txt.
    Returns (jENA row, samples)."""
    import os
    cc, cp, _ = read_solution(mesh_file, state_file)
    j, n = spectra(cc, cp, [the], n_v, device=device)
    write_spectrum(os.path.join(out_dir, f"jENA_{name}.txt"), j[0])
    return j[0], int(n[0])


def run_map(mesh_file, state_file, n_v=1000, out_dir=".", fix_high_energies=False, device=0):
    """main + Culc_ENA (:544-780): the map over map_angles(), written to
    <out_dir>/jENA_angle_test.txt.  Returns (angles, map)."""
    import os
    cc, cp, _ = read_solution(mesh_file, state_file)
    th = map_angles()
    j, _ = spectra(cc, cp, th, n_v, device=device)
    amap = angle_map(j, n_v, fix_high_energies)
    write_angle_map(os.path.join(out_dir, "jENA_angle_test.txt"), th, amap)
    return th, amap
