"""Host side of the ENA post-processor (set_amd/ena.py): the angles, speeds
and energies of SET_ENA.cpp, its linear_interpolate (:373-406), the angle
map's columns (:727-776) and the text layouts of its two output files
(:773-776, :934-952).  No GPU: the spectra here are given arrays."""
import numpy as np

from set_amd import ena, fcmhd, workloads


def test_map_angles_follow_reference_loop():
    """for (the = 0.008; the < pi - 0.008; the += pi/180): 180 angles, the
    last below pi - 0.008 and its successor not."""
    a = ena.map_angles()
    assert len(a) == 180 and a[0] == 0.008
    assert a[-1] < np.pi - 0.008 <= a[-1] + np.pi / 180.0
    np.testing.assert_allclose(np.diff(a), np.pi / 180.0, rtol=1e-11)


def test_speeds_and_energies():
    """v_i = 1 km/s + (i + 1/2) * 999 km/s / n_v, and E = 5.226e-3 (v in
    km/s)^2 eV (the hydrogen kinetic energy m_H v^2 / 2 at 1 km/s is
    5.2228e-3 eV: the reference's factor holds it to 0.1 %)."""
    v = ena.speeds(1000) / ena.KMS
    np.testing.assert_allclose(v, 1.0 + (np.arange(1000) + 0.5) * 0.999, rtol=1e-13)
    e = ena.energies_ev(1000)
    np.testing.assert_allclose(e, 5.226e-3 * v * v, rtol=1e-13)
    assert abs(0.5 * 1.6735575e-27 * 1e6 / 1.602176634e-19 / 5.226e-3 - 1.0) < 1e-3
    assert np.all(np.diff(e) > 0)


def test_linear_interpolate_clamps_and_interpolates():
    xs = np.array([1.0, 2.0, 4.0, 8.0])
    ys = np.array([10.0, 20.0, 0.0, 5.0])
    assert ena.linear_interpolate(0.5, xs, ys) == 10.0
    assert ena.linear_interpolate(1.0, xs, ys) == 10.0
    assert ena.linear_interpolate(8.0, xs, ys) == 5.0
    assert ena.linear_interpolate(9.0, xs, ys) == 5.0
    assert ena.linear_interpolate(2.0, xs, ys) == 20.0  # a knot takes its own value
    for x in np.linspace(1.0, 8.0, 57):
        assert abs(ena.linear_interpolate(x, xs, ys) - np.interp(x, xs, ys)) < 1e-13


def test_angle_map_columns():
    """Columns 1-8 at 15 ... 1821 eV; the reference evaluates columns 9-14 at
    1821 eV too (it passes E8 to each interpolation); fix_high_energies
    takes 450 ... 4290 eV.  A spectrum linear in energy makes each column
    its energy times the output factor."""
    n_v = 1000
    e = ena.energies_ev(n_v)
    j = np.vstack([e, 2.0 * e])
    m = ena.angle_map(j, n_v)
    want = np.array(ena.MAP_ENERGIES[:8] + (1821.0,) * 6) * ena.FLUX
    np.testing.assert_allclose(m[0], want, rtol=1e-12)
    np.testing.assert_allclose(m[1], 2.0 * want, rtol=1e-12)
    f = ena.angle_map(j, n_v, fix_high_energies=True)
    np.testing.assert_allclose(f[0], np.array(ena.MAP_ENERGIES) * ena.FLUX, rtol=1e-12)
    # beyond the last speed (1000 km/s: 5226 eV) the map clamps; 4290 eV is inside
    assert e[-1] > 4290.0


def test_spectrum_file_layout(tmp_path):
    """One line per speed: energy in keV, speed in km/s, intensity x flux,
    each with 6 significant digits as the C++ stream writes a double."""
    j = np.array([1.5e-70, 2.0, 0.0])
    p = tmp_path / "jENA_x.txt"
    ena.write_spectrum(p, j)
    lines = p.read_text().splitlines()
    assert len(lines) == 3
    v = ena.speeds(3) / ena.KMS
    first = lines[0].split()
    assert first[1] == "%g" % v[0] and float(first[1]) == float("%.6g" % v[0])
    assert first[2] == "1.01278e-62"          # 1.5e-70 * 6.75186e7
    assert lines[1].split()[2] == "1.35037e+08"
    assert lines[2].split()[2] == "0"
    assert float(first[0]) == float("%.6g" % (v[0] * v[0] * 5.226e-6))
    ena.write_spectrum(p, j, exact=True)
    assert float(p.read_text().split()[2]) == 1.5e-70 * ena.FLUX


def test_angle_map_file_layout(tmp_path):
    th = ena.map_angles()[:2]
    amap = np.arange(28.0).reshape(2, 14) * 1e-3
    p = tmp_path / "jENA_angle_test.txt"
    ena.write_angle_map(p, th, amap)
    rows = [r.split() for r in p.read_text().splitlines()]
    assert len(rows) == 2 and all(len(r) == 15 for r in rows)
    assert rows[0][0] == "0.008" and rows[1][0] == "%g" % th[1]
    assert rows[1][14] == "0.027"


def test_read_solution(tmp_path):
    """The mesh file gives the centres, the solver's output file the state."""
    fm, W = workloads.heliosphere_face_mesh(12, 6)
    fcmhd.write_input(tmp_path / "m.bin", fm, W, time=0.0)
    W2 = W.copy()
    W2[:, 0] *= 3.0
    fcmhd.save_storage(tmp_path / "s.bin", 1.25, W2)
    cc, cp, t = ena.read_solution(tmp_path / "m.bin", tmp_path / "s.bin")
    assert t == 1.25
    np.testing.assert_array_equal(cc, fm.cell_center)
    np.testing.assert_array_equal(cp, W2)


def test_triangulation_is_delaunay_with_consistent_neighbours():
    """ena.triangulate stands in for CGAL's Delaunay_triangulation_2
    (SET_ENA.cpp:242-257), which is unique for points in general position:
    every triangle counterclockwise, no centre strictly inside any
    triangle's circumcircle, and nb[t, k] the triangle across the edge
    opposite vertex k, sharing exactly that edge (-1 only on the hull)."""
    rng = np.random.default_rng(21)
    pts = np.c_[rng.uniform(-1, 1, 600), rng.uniform(0, 1, 600), np.zeros(600)]
    tri, nb = ena.triangulate(pts)
    p = pts[:, :2]
    a, b, c = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    orient = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1])
    assert np.all(orient > 0)
    # in-circle determinant of every point against every triangle (ccw: > 0 means inside)
    for lo in range(0, len(tri), 200):
        A, B, C = a[lo:lo + 200, None, :], b[lo:lo + 200, None, :], c[lo:lo + 200, None, :]
        d = p[None, :, :]
        m = np.stack([A - d, B - d, C - d], axis=2)  # (T, N, 3, 2)
        w = (m ** 2).sum(-1)
        det = (m[..., 0, 0] * (m[..., 1, 1] * w[..., 2] - w[..., 1] * m[..., 2, 1])
               - m[..., 0, 1] * (m[..., 1, 0] * w[..., 2] - w[..., 1] * m[..., 2, 0])
               + w[..., 0] * (m[..., 1, 0] * m[..., 2, 1] - m[..., 1, 1] * m[..., 2, 0]))
        assert det.max() < 1e-12
    for t in range(len(tri)):
        for k in range(3):
            edge = {tri[t, (k + 1) % 3], tri[t, (k + 2) % 3]}
            o = nb[t, k]
            if o < 0:
                continue
            assert edge <= set(tri[o]) and tri[t, k] not in tri[o]
            assert t in nb[o]
    # the hull edges are exactly the edges without a neighbour
    from scipy.spatial import ConvexHull
    assert (nb < 0).sum() == len(ConvexHull(p).vertices)
