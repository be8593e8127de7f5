"""Write golden vectors for chlld / minmod / linear from the reference itself.

Runs the reference's own routines (SET_Fortran_Cuda_MHD/Solvers.f90:108-150
and :1143-1859), compiled unmodified by oracle/ref_build.sh into
oracle/_ref/libref_chlld.so, on seeded random face states and on the special
cases that select each branch of chlld, and stores inputs and outputs in
tests/golden/chlld_golden.npz (plain arrays, no pickle).

    bash oracle/ref_build.sh && python tests/golden/make_golden.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "..", "oracle", "_ref", "libref_chlld.so")
OUT = os.path.join(HERE, "chlld_golden.npz")

D = ctypes.c_double
PD = ctypes.POINTER(D)


def load_ref():
    lib = ctypes.CDLL(LIB)
    lib.ref_chlld.argtypes = [ctypes.c_int, D, D, D, D, PD, PD, PD, PD, PD, PD, ctypes.c_int]
    lib.ref_chlld.restype = None
    lib.ref_minmod.argtypes = [D, D]
    lib.ref_minmod.restype = D
    lib.ref_linear.argtypes = [D] * 7
    lib.ref_linear.restype = D
    return lib


def ref_chlld(lib, n_state, nrm, q1, q2):
    a = (D * 8)(*q1)
    b = (D * 8)(*q2)
    out = (D * 8)()
    dl, dp, dc = D(), D(), D()
    lib.ref_chlld(int(n_state), float(nrm[0]), float(nrm[1]), float(nrm[2]), 0.0, a, b,
                  ctypes.byref(dl), ctypes.byref(dp), ctypes.byref(dc), out, 0)
    return np.array(list(out) + [dl.value, dp.value, dc.value])


def prim(rng, rho=(0.05, 2.0), vel=1.5, p=(0.05, 2.0), b=1.0):
    return np.concatenate([[rng.uniform(*rho)], rng.normal(scale=vel, size=3),
                           [rng.uniform(*p)], rng.normal(scale=b, size=3)])


def cases(rng):
    """(normal, q1, q2) triples: random plus branch-selecting special cases."""
    out = []
    axes = [np.array(v, float) for v in
            ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1])]
    for i in range(160):  # random states, random and axis normals
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        if i % 3 == 0:
            n = axes[i % 6]
        out.append((n, prim(rng), prim(rng)))
    for i in range(12):  # identical states (consistency F(U,U))
        q = prim(rng)
        out.append((axes[i % 6], q, q.copy()))
    for i in range(12):  # B = 0: the d <= eps basis branch, hydro HLLC
        q1, q2 = prim(rng, b=0.0), prim(rng, b=0.0)
        out.append((axes[i % 6] if i % 2 else rng.normal(size=3) / np.sqrt(3), q1, q2))
    for i in range(12):  # B along the normal only (d <= eps with B != 0)
        n = axes[i % 6]
        q1, q2 = prim(rng), prim(rng)
        q1[5:8] = n * 0.7
        q2[5:8] = n * 0.7
        out.append((n, q1, q2))
    for i in range(12):  # tangential B only: |Bn| < epsb switch in HLLC/HLLD
        n = axes[i % 6]
        q1, q2 = prim(rng), prim(rng)
        q1[5:8] -= n * np.dot(n, q1[5:8])
        q2[5:8] -= n * np.dot(n, q2[5:8])
        out.append((n, q1, q2))
    for i in range(12):  # supersonic to +n (SL > 0) and to -n (SR < 0)
        n = axes[i % 6]
        q1, q2 = prim(rng, vel=0.2, b=0.1), prim(rng, vel=0.2, b=0.1)
        s = 8.0 if i % 2 == 0 else -8.0
        q1[1:4] += s * n
        q2[1:4] += s * n
        out.append((n, q1, q2))
    for i in range(12):  # mirror-symmetric pair: SM = 0 ties
        n = axes[i % 6]
        q1 = prim(rng)
        q2 = q1.copy()
        vn = np.dot(n, q1[1:4])
        q2[1:4] = q1[1:4] - 2 * vn * n
        out.append((n, q1, q2))
    # the wind / inflow states of the Cartesian heliosphere workload
    wind = np.array([100.0, 10.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0])
    inflow = np.array([0.1, -2.5, 0.0, 0.0, 0.1,
                       0.3 * np.cos(np.pi / 6), 0.3 * np.sin(np.pi / 6), 0.0])
    for n in axes:
        out.append((n, wind, inflow))
        out.append((n, inflow, wind))
        out.append((n, inflow, inflow * np.array([1.01, 1, 1, 1, 1, 1, 1, 1])))
    return out


def main():
    lib = load_ref()
    rng = np.random.default_rng(20261203)
    cs = cases(rng)
    normals = np.array([c[0] for c in cs])
    q1 = np.array([c[1] for c in cs])
    q2 = np.array([c[2] for c in cs])
    flux = np.zeros((4, len(cs), 11))
    for ns in range(4):
        for k in range(len(cs)):
            flux[ns, k] = ref_chlld(lib, ns, normals[k], q1[k], q2[k])
    # minmod / linear tables
    vals = np.array([-2.0, -0.5, -0.0, 0.0, 0.25, 3.0, 1e-300, -1e300])
    mm_x, mm_y = np.meshgrid(vals, vals)
    mm = np.array([[lib.ref_minmod(float(a), float(b)) for a, b in zip(r1, r2)]
                   for r1, r2 in zip(mm_x, mm_y)])
    lin_in = rng.normal(size=(200, 7))
    lin_in[:, 0] = -np.abs(lin_in[:, 0]) - 1.0  # x1 < x2 < x3 as in the face pass
    lin_in[:, 2] = -np.abs(lin_in[:, 2]) * 0.5
    lin_in[:, 4] = np.abs(lin_in[:, 4]) + 0.1
    lin_in[:, 6] = 0.0
    lin_in[::5, 3] = lin_in[::5, 1]  # flat left slope -> minmod 0
    lin = np.array([lib.ref_linear(*map(float, row)) for row in lin_in])
    np.savez_compressed(OUT, normals=normals, q1=q1, q2=q2, flux=flux,
                        mm_x=mm_x, mm_y=mm_y, mm=mm, lin_in=lin_in, lin=lin)
    print(f"wrote {OUT}: {len(cs)} face cases x 4 solvers, {mm.size} minmod, {len(lin)} linear")


if __name__ == "__main__":
    main()
