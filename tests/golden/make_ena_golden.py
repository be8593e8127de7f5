"""Writes tests/golden/ena_golden.npz: input/output vectors of the
reference's own ENA rate functions (SET_ENA/SET_ENA.cpp:341-370, 426-542:
Velosity_1, maxwell_distribution, int_1_f1/f2/f3, int_1), compiled from the
reference's source where it lies by oracle/ref_build.sh into
oracle/_ref/libref_ena.so.  Run here, where the reference is mounted:

    bash oracle/ref_build.sh && python tests/golden/make_ena_golden.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "..", "oracle", "_ref", "libref_ena.so")


def main():
    L = ctypes.CDLL(LIB)
    D = ctypes.c_double
    for f in ("ref_ena_velocity1", "ref_ena_int1"):
        getattr(L, f).argtypes = [D, D]
        getattr(L, f).restype = D
    L.ref_ena_f.argtypes = [ctypes.c_int, D]
    L.ref_ena_f.restype = D
    L.ref_ena_maxwell.argtypes = [D] * 8
    L.ref_ena_maxwell.restype = D
    rng = np.random.default_rng(2025)
    # every branch of the fits (x <= 1, 3, 5, 7) and their joins
    fx = np.r_[np.linspace(0.0, 7.0, 141), rng.uniform(0.0, 7.0, 300), [1.0, 3.0, 5.0, 7.0]]
    f = np.array([[L.ref_ena_f(w, x) for x in fx] for w in (1, 2, 3)])
    cp = rng.uniform(0.05, 5.0, 400)
    xu = rng.uniform(0.0, 7.0, 400) * cp
    int1 = np.array([L.ref_ena_int1(a, b) for a, b in zip(xu, cp)])
    # both branches of Velosity_1 (u < 1e-5 series)
    vu = np.r_[rng.uniform(0.0, 1e-5, 100), rng.uniform(1e-5, 30.0, 300)]
    vcp = rng.uniform(0.05, 5.0, 400)
    vel = np.array([L.ref_ena_velocity1(a, b) for a, b in zip(vu, vcp)])
    mx_in = np.c_[rng.uniform(0.01, 2, 300), rng.uniform(0.1, 3, 300), rng.normal(size=(300, 6))]
    mx = np.array([L.ref_ena_maxwell(*map(float, r)) for r in mx_in])
    np.savez_compressed(os.path.join(HERE, "ena_golden.npz"), f_x=fx, f=f, int1_x=xu, int1_cp=cp,
                        int1=int1, vel_u=vu, vel_cp=vcp, vel=vel, maxwell_in=mx_in, maxwell=mx)


if __name__ == "__main__":
    main()
