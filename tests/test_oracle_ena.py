"""The oracle's ENA line-of-sight integration (oracle/ena.c, restating
SET_ENA/SET_ENA.cpp:341-956) against the reference's own rate functions and
against what the mathematics fixes.

Pins:
  * golden vectors of the reference's Velosity_1, maxwell_distribution,
    int_1_f1/f2/f3 and int_1, compiled from SET_ENA.cpp where it lies
    (oracle/ref_build.sh; fixture by tests/golden/make_ena_golden.py) --
    bitwise; and fresh random inputs when oracle/_ref/libref_ena.so is built;
  * each fit against the integral it approximates (quadrature): F1(y) =
    int |z| exp(-|z - y|^2) d^3z, F2 and F3 with ln|z| and ln^2|z|;
    Velosity_1 against the Maxwellian mean relative speed;
  * the line of sight: the sample count a straight march through a
    rectangle gives, the spectrum of a uniform state in closed form,
    linearity in the neutral density, a spectrum that peaks at the speed of
    plasma streaming toward the observer, and mirror symmetry.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from set_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ena_golden.npz")
REF = os.path.join(os.path.dirname(oracle.__file__), "_ref", "libref_ena.so")
AE, KMS = 0.00313166, 0.0963358
A2, KN, SQRTPI = 0.1307345665, 0.4326569, 1.77245385


def test_rate_functions_match_reference_golden():
    g = np.load(GOLD)
    for k in (1, 2, 3):
        got = np.array([oracle.ena_f(k, x) for x in g["f_x"]])
        np.testing.assert_array_equal(got, g["f"][k - 1])
    np.testing.assert_array_equal([oracle.ena_int1(x, c) for x, c in zip(g["int1_x"], g["int1_cp"])], g["int1"])
    np.testing.assert_array_equal([oracle.ena_velocity1(u, c) for u, c in zip(g["vel_u"], g["vel_cp"])], g["vel"])
    np.testing.assert_array_equal([oracle.ena_maxwell(*r) for r in g["maxwell_in"]], g["maxwell"])


@pytest.mark.skipif(not os.path.exists(REF), reason="reference ENA build absent")
def test_rate_functions_match_reference_build_random():
    L = ctypes.CDLL(REF)
    D = ctypes.c_double
    for f in ("ref_ena_velocity1", "ref_ena_int1"):
        getattr(L, f).argtypes = [D, D]
        getattr(L, f).restype = D
    L.ref_ena_f.argtypes = [ctypes.c_int, D]
    L.ref_ena_f.restype = D
    rng = np.random.default_rng(9)
    for _ in range(3000):
        x, cp, u = rng.uniform(0, 7), rng.uniform(0.05, 5), rng.uniform(0, 30)
        k = int(rng.integers(1, 4))
        assert oracle.ena_f(k, x) == L.ref_ena_f(k, x)
        assert oracle.ena_int1(x * cp, cp) == L.ref_ena_int1(x * cp, cp)
        assert oracle.ena_velocity1(u, cp) == L.ref_ena_velocity1(u, cp)


def fit_integral(k, y):
    """int |z| ln^(k-1)|z| exp(-|z - y|^2) d^3z, by quadrature over |z| after
    the angular integral."""
    from scipy.integrate import quad

    def g(z):
        w = [1.0, np.log(z), np.log(z) ** 2][k - 1]
        ang = 4 * np.pi * np.exp(-z * z) if y == 0 else \
            np.pi / (y * z) * (np.exp(-(z - y) ** 2) - np.exp(-(z + y) ** 2))
        return z ** 3 * w * ang
    return quad(g, 1e-300, y + 12, limit=500, points=[y] if y > 0 else None)[0]


@pytest.mark.parametrize("k", [1, 2, 3])
def test_fits_approximate_their_integrals(k):
    for y in np.r_[0.0, np.linspace(0.25, 7.0, 28), [1.0, 3.0, 5.0]]:
        want = fit_integral(k, y)
        # the reference's fits hold their integrals to 1e-5 (f1), 2e-6 (f2), 4e-5 (f3)
        assert oracle.ena_f(k, y) == pytest.approx(want, rel=5e-5), (k, y)


def test_velocity1_is_mean_relative_speed():
    """<|u - w|> over a Maxwellian of thermal speed cp: cp (exp(-x^2)/sqrt(pi)
    + (x + 1/(2x)) erf(x)), x = u/cp, from the quadrature of fit 1; the
    reference's sqrt(pi) 1.77245385 is truncated at 9 digits."""
    for cp in (0.3, 1.0, 2.5):
        for x in (1e-8, 0.3, 2.0, 7.5, 12.0):
            want = cp * fit_integral(1, x) / np.pi ** 1.5
            assert oracle.ena_velocity1(x * cp, cp) == pytest.approx(want, rel=1e-8)


def rect_mesh(nx, ny, x0=-1.0, x1=1.0, y0=-1.0, y1=1.0):
    g = mesh.Grid(nx, ny, 1, x0, x1, y0, y1)
    cc = g.cell_centers()
    # two triangles per quad of neighbouring centres, counterclockwise
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="xy")
    a = (i + nx * j).ravel()
    tri = np.concatenate([np.c_[a, a + 1, a + 1 + nx], np.c_[a, a + 1 + nx, a + nx]]).astype(np.int32)
    return g, cc, tri


def uniform_par(n, rho=0.2, u=-1.0, v=0.3, p=0.05, nh=0.1, uh=-2.0, vh=0.1, th=0.4):
    W = np.zeros((n, 16))
    W[:, [0, 1, 2, 4]] = [rho, u, v, p]
    W[:, 8:12] = [nh, uh, vh, th]
    W[:, 12:16] = [nh * 0.5, uh * 0.7, vh, th * 2.0]
    return W


def test_uniform_state_closed_form():
    """Uniform plasma and neutrals: every rate is the same at every sample but
    photoionisation (770.291 (AU/r)^2); S after sample k is exp of minus the
    summed rates times ds/v, and j = nu_H f(-v l) v ds (S_1 + ... + S_n).  The
    samples are the points 5 AU + ds/2 + k ds along the line inside the hull
    of the cell centres."""
    g, cc, tri = rect_mesh(40, 20)
    W = uniform_par(len(cc))
    the = np.deg2rad(60.0)
    n_v = 16
    j, n = oracle.ena_los(cc, W, tri, the, n_v)
    lx, ly = np.cos(the), np.sin(the)
    ds = AE / 30
    s = 5 * AE + ds / 2 + ds * np.arange(100000)
    pts = np.c_[s * lx, s * ly]
    lo, hi = cc[:, :2].min(0), cc[:, :2].max(0)
    inside = np.all((pts >= lo - 1e-13) & (pts <= hi + 1e-13), axis=1)
    n_expect = int(np.argmin(inside))
    assert inside[:n_expect].all() and n == n_expect > 100
    r = s[:n]
    v = (1 + (np.arange(n_v) + 0.5) * (999 / n_v)) * KMS

    def rate(uu, rho, cp):
        return np.array([rho * oracle.ena_int1(x, cp) / KN if x / cp <= 7 else
                         rho * oracle.ena_velocity1(x, cp) * (1 - A2 * np.log(oracle.ena_velocity1(x, cp))) ** 2 / KN
                         for x in uu])

    wx, wy = -v * lx, -v * ly
    rho, u, vv, p = 0.2, -1.0, 0.3, 0.05
    cpp = np.sqrt(p / rho)
    nu_ex = rate(np.hypot(wx - u, wy - vv), rho, cpp)
    nu_h = (rate(np.hypot(wx + 2.0, wy - 0.1), 0.3, np.sqrt(0.2)) +
            rate(np.hypot(wx + 1.4, wy - 0.1), 0.15, np.sqrt(0.4)))
    f = rho / (np.pi * np.sqrt(np.pi) * cpp ** 3) * np.exp(-((wx - u) ** 2 + (wy - vv) ** 2) / cpp ** 2)
    nu_ph = 770.291 * (AE / r) ** 2
    expo = np.cumsum(nu_ex[None, :] + nu_ph[:, None], axis=0) * ds / v[None, :]
    want = nu_h * f * v * ds * np.exp(-expo).sum(0)
    np.testing.assert_allclose(j, want, rtol=1e-10)


def test_linear_in_neutral_density():
    """Doubling both neutral densities doubles the ENA production and leaves
    the survival (set by the plasma and photoionisation) alone: j doubles
    bit for bit."""
    g, cc, tri = rect_mesh(30, 15)
    rng = np.random.default_rng(3)
    W = uniform_par(len(cc), p=2.0)  # thermal speed ~33 km/s
    W[:, 0] *= rng.uniform(0.8, 1.2, len(cc))
    W[:, 8] *= rng.uniform(0.5, 1.5, len(cc))
    j1, n = oracle.ena_los(cc, W, tri, np.deg2rad(75.0), 24)
    assert n > 100 and np.all(j1[:3] > 0)
    W2 = W.copy()
    W2[:, [8, 12]] *= 2.0
    j2, _ = oracle.ena_los(cc, W2, tri, np.deg2rad(75.0), 24)
    np.testing.assert_array_equal(j2, 2.0 * j1)


def test_spectrum_peaks_at_inflow_toward_observer():
    """Plasma streaming toward the observer at speed V along the line of sight
    makes ENAs that arrive at about speed V: the spectrum's maximum sits at
    the speed bin of V (a mis-signed particle velocity puts it nowhere)."""
    g, cc, tri = rect_mesh(30, 15)
    the = np.deg2rad(50.0)
    V = 300 * KMS
    W = uniform_par(len(cc), rho=0.05, u=-V * np.cos(the), v=-V * np.sin(the), p=0.05 * (20 * KMS) ** 2)
    j, n = oracle.ena_los(cc, W, tri, the, 200)
    v = (1 + (np.arange(200) + 0.5) * (999 / 200)) * KMS
    assert n > 0 and abs(v[np.argmax(j)] - V) < 10 * KMS


def test_mirror_symmetry():
    """A state and a triangulation symmetric under x -> -x, seen along angles
    the and pi - the, give the same spectrum (to rounding: the samples are
    mirrored positions)."""
    nx, ny = 41, 20
    g = mesh.Grid(nx, ny, 1, -1.0, 1.0, -1.0, 1.0)
    cc = g.cell_centers()
    tris = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            a = i + nx * j
            if i < (nx - 1) // 2:  # left half: diagonal lower-left to upper-right
                tris += [[a, a + 1, a + 1 + nx], [a, a + 1 + nx, a + nx]]
            else:  # right half: its mirror image
                tris += [[a, a + 1, a + nx], [a + 1, a + 1 + nx, a + nx]]
    tri = np.array(tris, np.int32)
    rng = np.random.default_rng(4)
    W = uniform_par(len(cc), p=2.0)
    half = rng.uniform(0.8, 1.2, (ny, nx // 2 + 1))
    sym = np.concatenate([half[:, :0:-1], half], axis=1).ravel()  # column c equals column nx-1-c
    W[:, 0] *= sym
    W[:, 8] *= sym
    W[:, [1, 9, 13]] = 0.0  # no x-velocity: the mirror image is the same state
    a, na = oracle.ena_los(cc, W, tri, np.deg2rad(70.0), 32)
    b, nb = oracle.ena_los(cc, W, tri, np.pi - np.deg2rad(70.0), 32)
    assert na == nb > 100 and np.all(a[:3] > 0)
    np.testing.assert_allclose(a, b, rtol=1e-9)
