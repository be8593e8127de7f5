"""Text outputs in the formats of SET's Fortran writers (FCMHD.f90:6-119):
the Fortran edit descriptors they use, the format reversion of Print, and
the dimensional scale factors."""
import numpy as np

from set_amd import tecplot


def test_fortran_es_descriptor():
    assert tecplot._es(1.0, 15, 7) == "  1.0000000E+00"
    assert tecplot._es(-0.000123456789, 15, 7) == " -1.2345679E-04"
    assert tecplot._es(146.262, 12, 5) == " 1.46262E+02"
    assert len(tecplot._es(3.0, 23, 15)) == 23


def cells():
    cc = np.array([[0.5, 0.25, 0.0], [1.0, 2.0, 0.0]])
    cp = np.zeros((2, 16))
    cp[0, :5] = [2.0, 3.0, 4.0, 0.0, 5.0]
    cp[1, :5] = [1.0, -1.0, 0.5, 0.0, 0.1]
    return cc, cp


def test_state_file_reversion_and_units(tmp_path):
    cc, cp = cells()
    p = tmp_path / "state.txt"
    tecplot.write_state(p, cc, cp)
    lines = p.read_text().splitlines()
    assert lines[0] == 'TITLE = "HP"  VARIABLES = x, y, rho, u, v, p, |V|'
    # '(6(ES15.7, 2X))' with 7 items: six on a line, |V| on the next
    assert len(lines) == 1 + 2 * 2
    first = [float(v) for v in lines[1].split()]
    assert len(first) == 6 and len(lines[2].split()) == 1
    u, v = 3.0 * 10.3565, 4.0 * 10.3565
    np.testing.assert_allclose(first, [0.5 * 319.319, 0.25 * 319.319, 2.0 * 0.073, u, v, 5.0 * 1.30962],
                               rtol=1e-7)
    np.testing.assert_allclose(float(lines[2]), np.hypot(u, v), rtol=1e-7)


def test_density_file(tmp_path):
    cc, cp = cells()
    p = tmp_path / "dens.txt"
    tecplot.write_density(p, cc, cp, 0.5)
    lines = p.read_text().splitlines()
    assert lines[:2] == ['TITLE = "HP"', "VARIABLES = x, y, rho"]
    assert lines[2] == 'ZONE T="HP", STRANDID=1, SOLUTIONTIME=' + tecplot._es(0.5 * 146.262, 12, 5)
    vals = np.array([[float(v) for v in l.split()] for l in lines[3:]])
    np.testing.assert_allclose(vals, [[0.5 * 319.319, 0.25 * 319.319, 2.0 * 0.073],
                                      [319.319, 2.0 * 319.319, 0.073]], rtol=1e-7)


def test_probe_series_appends(tmp_path):
    p = tmp_path / "probes.txt"
    tecplot.append_probes(p, 1.0, np.arange(1, 18))
    tecplot.append_probes(p, 2.0, np.arange(1, 18))
    lines = p.read_text().splitlines()
    assert len(lines) == 2 and all(len(l) == 18 * 23 for l in lines)
    vals = [float(v) for v in lines[1].split()]
    # the reference's factors are single-precision literals promoted to double
    assert tecplot.SCALE_RHO == float(np.float32(0.073)) != 0.073
    np.testing.assert_allclose(vals, [2.0 * tecplot.SCALE_T] + [k * tecplot.SCALE_RHO for k in range(1, 18)],
                               rtol=1e-15)
