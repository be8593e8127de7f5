"""The CPU oracle against the reference's own chlld and against what the
mathematics fixes.

Pins, in order of strength:
  * golden vectors written by the reference's chlld / minmod / linear
    (Solvers.f90, compiled unmodified by oracle/ref_build.sh; fixture made
    by tests/golden/make_golden.py) -- bitwise;
  * when oracle/_ref/libref_chlld.so is present, the same on fresh random
    states;
  * closed forms: consistency F(U,U) = F(U), supersonic upwinding, HLL
    against the textbook formula restated here in numpy, contact
    resolution by HLLC/HLLD, mirror symmetry, rotation invariance.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "chlld_golden.npz")
SPI4 = np.sqrt(12.56637061435917295384)
GAMMA = 5.0 / 3.0


def test_chlld_matches_reference_golden_bitwise():
    g = np.load(GOLD)
    n = g["normals"]
    for ns in range(4):
        for k in range(len(n)):
            out, dl, dp, dc, st = oracle.chlld(ns, n[k], g["q1"][k], g["q2"][k])
            got = np.concatenate([out, [dl, dp, dc]])
            assert st == 0
            np.testing.assert_array_equal(got, g["flux"][ns, k], err_msg=f"n_state={ns} case={k}")


def test_minmod_linear_match_reference_golden():
    g = np.load(GOLD)
    mm = np.array([[oracle.minmod(a, b) for a, b in zip(r1, r2)]
                   for r1, r2 in zip(g["mm_x"], g["mm_y"])])
    np.testing.assert_array_equal(mm, g["mm"])
    lin = np.array([oracle.linear(*row) for row in g["lin_in"]])
    np.testing.assert_array_equal(lin, g["lin"])


@pytest.mark.skipif(not os.path.exists(oracle.REF_PATH), reason="reference build absent")
def test_chlld_matches_reference_build_random():
    lib = ctypes.CDLL(oracle.REF_PATH)
    D, PD = ctypes.c_double, ctypes.POINTER(ctypes.c_double)
    lib.ref_chlld.argtypes = [ctypes.c_int, D, D, D, D, PD, PD, PD, PD, PD, PD, ctypes.c_int]
    rng = np.random.default_rng(7)
    for it in range(3000):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        q1 = np.r_[rng.uniform(0.05, 3), rng.normal(size=3), rng.uniform(0.05, 3), rng.normal(size=3)]
        q2 = np.r_[rng.uniform(0.05, 3), rng.normal(size=3), rng.uniform(0.05, 3), rng.normal(size=3)]
        ns = it % 4
        a, b, o = (D * 8)(*q1), (D * 8)(*q2), (D * 8)()
        dl, dp, dc = D(), D(), D()
        lib.ref_chlld(ns, *map(float, nrm), 0.0, a, b, ctypes.byref(dl), ctypes.byref(dp),
                      ctypes.byref(dc), o, 0)
        ref = np.r_[list(o), dl.value, dp.value, dc.value]
        out, l, p, c, st = oracle.chlld(ns, nrm, q1, q2)
        np.testing.assert_array_equal(np.r_[out, l, p, c], ref)


# ---------------------------------------------------------------------------
# closed forms (independent numpy restatements)

def phys_flux(q, n):
    """Ideal-MHD flux through unit normal n of primitive state q (SET's
    stored B: b = B/sqrt(4 pi)); returns the same 8 components chlld
    returns, normal-B component 0."""
    r, v, p, B = q[0], q[1:4], q[4], q[5:8]
    b = B / SPI4
    vn, bn = v @ n, b @ n
    pt = p + 0.5 * b @ b
    e = p / (GAMMA - 1) + 0.5 * r * v @ v + 0.5 * b @ b
    F = np.zeros(8)
    F[0] = r * vn
    F[1:4] = r * vn * v + pt * n - bn * b
    F[4] = (e + pt) * vn - bn * (v @ b)
    F[5:8] = (vn * b - bn * v) * SPI4
    return F


def fast_speed(q, n):
    r, p, b = q[0], q[4], q[5:8] / SPI4
    c2 = GAMMA * p / r
    a2 = b @ b / r
    an2 = (b @ n) ** 2 / r
    return np.sqrt(0.5 * (c2 + a2 + np.sqrt((c2 + a2) ** 2 - 4 * c2 * an2)))


def cons(q):
    r, v, p, B = q[0], q[1:4], q[4], q[5:8]
    b = B / SPI4
    return np.r_[r, r * v, p / (GAMMA - 1) + 0.5 * r * v @ v + 0.5 * b @ b, B]


def rand_state(rng, b=0.8):
    return np.r_[rng.uniform(0.2, 2), rng.normal(size=3) * 0.5, rng.uniform(0.2, 2),
                 rng.normal(size=3) * b]


@pytest.mark.parametrize("ns", [1, 2, 3])
def test_consistency_identical_states(ns):
    rng = np.random.default_rng(11)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        q = rand_state(rng)
        out = oracle.chlld(ns, n, q, q)[0]
        ref = phys_flux(q, n)
        # normal component of the B flux is the Bn-jump term, 0 for equal states
        np.testing.assert_allclose(out, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("ns", [1, 2, 3])
def test_supersonic_upwind(ns):
    rng = np.random.default_rng(12)
    for sgn in (1.0, -1.0):
        for _ in range(50):
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            q1, q2 = rand_state(rng, 0.2), rand_state(rng, 0.2)
            q2[5:8] = q1[5:8]  # equal normal B: no normal-B jump term
            q1[1:4] += sgn * 12.0 * n
            q2[1:4] += sgn * 12.0 * n
            out, sl, sr, sm, st = oracle.chlld(ns, n, q1, q2)
            assert (sl > 0) if sgn > 0 else (sr < 0)
            ref = phys_flux(q1 if sgn > 0 else q2, n)
            np.testing.assert_allclose(out, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


def test_hll_textbook():
    """n_state=1 against F = (SR FL - SL FR + SL SR (UR - UL)) / (SR - SL)
    with Davis speeds SL = min(vnL - cfL, vnR - cfR), SR = max(...)."""
    rng = np.random.default_rng(13)
    for _ in range(300):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        q1, q2 = rand_state(rng), rand_state(rng)
        vl, vr = q1[1:4] @ n, q2[1:4] @ n
        cl, cr = fast_speed(q1, n), fast_speed(q2, n)
        SL, SR = min(vl - cl, vr - cr), max(vl + cl, vr + cr)
        FL, FR = phys_flux(q1, n), phys_flux(q2, n)
        UL, UR = cons(q1), cons(q2)
        if SL > 0:
            ref = FL
        elif SR < 0:
            ref = FR
        else:
            ref = (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)
        out, sl, sr, sm, st = oracle.chlld(1, n, q1, q2)
        assert sl == pytest.approx(SL, rel=1e-12) and sr == pytest.approx(SR, rel=1e-12)
        np.testing.assert_allclose(out, ref, rtol=1e-10, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("ns", [2, 3])
def test_contact_resolved_exactly(ns):
    """An isolated contact (density jump, equal p and v, vn = 0, Bn = 0)
    carries no mass flux in HLLC/HLLD; HLL smears it."""
    n = np.array([1.0, 0.0, 0.0])
    q1 = np.array([1.0, 0.0, 0.3, -0.2, 1.0, 0.0, 0.4, 0.1])
    q2 = np.array([0.125, 0.0, 0.3, -0.2, 1.0, 0.0, 0.4, 0.1])
    out = oracle.chlld(ns, n, q1, q2)[0]
    assert abs(out[0]) < 1e-14
    assert abs(oracle.chlld(1, n, q1, q2)[0][0]) > 1e-3


@pytest.mark.parametrize("ns", [0, 1, 2, 3])
def test_mirror_symmetry_zero_mass_flux(ns):
    rng = np.random.default_rng(14)
    n = np.array([0.0, 0.0, 1.0])
    for _ in range(50):
        q1 = rand_state(rng)
        q2 = q1.copy()
        q2[3] = -q1[3]  # mirror the normal velocity
        q2[7] = -q1[7]  # and the normal field
        out = oracle.chlld(ns, n, q1, q2)[0]
        assert abs(out[0]) < 1e-13 * (1 + abs(q1[0] * q1[3]))


@pytest.mark.parametrize("ns", [1, 2, 3])
def test_rotation_invariance(ns):
    rng = np.random.default_rng(15)
    for _ in range(100):
        A, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        q1, q2 = rand_state(rng), rand_state(rng)
        r1, r2 = q1.copy(), q2.copy()
        for q, r in ((q1, r1), (q2, r2)):
            r[1:4] = A @ q[1:4]
            r[5:8] = A @ q[5:8]
        f = oracle.chlld(ns, n, q1, q2)[0]
        g = oracle.chlld(ns, A @ n, r1, r2)[0]
        fr = f.copy()
        fr[1:4] = A @ f[1:4]
        fr[5:8] = A @ f[5:8]
        np.testing.assert_allclose(g, fr, rtol=1e-9, atol=1e-10 * np.abs(f).max())


def test_hlld_equals_hllc_without_field():
    rng = np.random.default_rng(16)
    for _ in range(100):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        q1, q2 = rand_state(rng, 0.0), rand_state(rng, 0.0)
        a = oracle.chlld(2, n, q1, q2)[0]
        b = oracle.chlld(3, n, q1, q2)[0]
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-12 * np.abs(a).max())


def test_minmod_closed_form():
    vals = [-3.0, -1.0, -0.25, 0.0, 0.5, 2.0]
    for x in vals:
        for y in vals:
            ref = 0.0 if x * y <= 0 else np.sign(x) * min(abs(x), abs(y))
            assert oracle.minmod(x, y) == ref


def test_linear_is_minmod_muscl_on_uniform_grid():
    """On a uniform grid linear(-1.5h, ...) is the minmod-limited face value."""
    rng = np.random.default_rng(17)
    h = 0.01
    for _ in range(200):
        ta, tb, tc = rng.normal(size=3)
        got = oracle.linear(-1.5 * h, ta, -0.5 * h, tb, 0.5 * h, tc, 0.0)
        sa, sb = (tb - ta) / h, (tc - tb) / h
        s = 0.0 if sa * sb <= 0 else np.sign(sa) * min(abs(sa), abs(sb))
        assert got == pytest.approx(tb + 0.5 * h * s, rel=1e-12, abs=1e-15)


def test_source_mk_closed_form():
    """Source_MK (cuf_kernel.cuf:215-288) against a numpy restatement."""
    rng = np.random.default_rng(18)
    for _ in range(50):
        cp = np.r_[rng.uniform(0.1, 2), rng.normal(size=3), rng.uniform(0.1, 2), np.zeros(3),
                   rng.uniform(0.01, 1), rng.normal(size=2), rng.uniform(0.1, 2),
                   rng.uniform(0.01, 1), rng.normal(size=2), rng.uniform(0.1, 2)]
        rho, u, v, p = cp[0], cp[1], cp[2], cp[4]
        S = np.zeros(3)
        for base in (8, 12):
            rh, uh, vh, th = cp[base:base + 4]
            ph = 0.5 * th * rh
            um = np.sqrt((u - uh) ** 2 + (v - vh) ** 2 + 64 / (9 * np.pi) * (p / rho + 2 * ph / rh))
            uu = np.sqrt((u - uh) ** 2 + (v - vh) ** 2 + 4 / np.pi * (p / rho + 2 * ph / rh))
            sig = (1 - np.float32(0.1307345665) * np.log(um)) ** 2
            nu = rho * rh * um * sig
            k = np.float32(3.0) / np.float32(0.4326569)
            S += k * nu * np.array([uh - u, vh - v,
                                    (uh ** 2 + vh ** 2 - u ** 2 - v ** 2) / 2
                                    + uu / um * (2 * ph / rh - p / rho)])
        np.testing.assert_allclose(oracle.source_mk(cp), S, rtol=1e-12)


def test_source_mk_vanishes_for_comoving_equal_temperature():
    """Neutrals moving with the plasma at its temperature (2 pH/rhoH = p/rho,
    pH = T rhoH / 2) exchange no momentum or energy."""
    rng = np.random.default_rng(19)
    for _ in range(50):
        rho, u, v, p = rng.uniform(0.1, 2), rng.normal(), rng.normal(), rng.uniform(0.1, 2)
        T = p / rho
        cp = np.r_[rho, u, v, 0.0, p, 0, 0, 0, rng.uniform(0.01, 1), u, v, T,
                   rng.uniform(0.01, 1), u, v, T]
        S = oracle.source_mk(cp)
        scale = rho * (1 + u * u + v * v + T)
        assert np.all(np.abs(S) <= 1e-12 * scale), S


def test_source_mk_linear_in_neutral_density():
    rng = np.random.default_rng(20)
    for _ in range(50):
        cp = np.r_[rng.uniform(0.1, 2), rng.normal(size=3), rng.uniform(0.1, 2), np.zeros(3),
                   rng.uniform(0.01, 1), rng.normal(size=2), rng.uniform(0.1, 2),
                   rng.uniform(0.01, 1), rng.normal(size=2), rng.uniform(0.1, 2)]
        c2 = cp.copy()
        c2[[8, 12]] *= 2.0
        np.testing.assert_allclose(oracle.source_mk(c2), 2.0 * oracle.source_mk(cp), rtol=1e-13)
