// ref_ena_bind.cpp -- C binding for the reference's own ENA rate functions
// (Velosity_1, maxwell_distribution, int_1_f1/f2/f3, int_1), compiled from
// SET_ENA/SET_ENA.cpp where they lie (see ref_build.sh, which cuts those
// routines and the macros they use into _ref/ref_ena_cut.inc).  Test
// infrastructure: it only makes the unmodified reference routines callable
// from the oracle's tests.
#include <cmath>
#include <iostream>
using namespace std;
#include "ref_ena_cut.inc"

extern "C" {
double ref_ena_velocity1(double u, double cp) { return Velosity_1(u, cp); }
double ref_ena_f(int which, double x) {
  return which == 1 ? int_1_f1(x) : which == 2 ? int_1_f2(x) : int_1_f3(x);
}
double ref_ena_int1(double x, double cp) { return int_1(x, cp); }
double ref_ena_maxwell(double rho, double cp, double u, double v, double w, double vx, double vy,
                       double vz) {
  return maxwell_distribution(rho, cp, u, v, w, vx, vy, vz);
}
}
