"""The C ABI library loads and exports every entry point include/setmhd.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from set_amd import solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "setmhd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(setmhd_[a-z_0-9]+)\s*\(", text)))


def test_header_lists_the_binding_exports():
    assert declared() == sorted(solver.EXPORTS)


@pytest.mark.skipif(not os.path.exists(solver.LIB_PATH), reason="libsetmhd.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(solver.LIB_PATH)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(solver.LIB_PATH), reason="libsetmhd.so not built")
def test_config_defaults_without_device():
    ref = solver.reference_config()
    assert (ref.riemann, ref.recon, ref.time_scheme) == (2, 1, 1)
    assert ref.cfl == float(__import__("numpy").float32(0.2))
    car = solver.cartesian_config(cfl=0.4, r_hold=0.1)
    assert (car.riemann, car.recon, car.time_scheme, car.flags) == (3, 0, 0, 0)
    assert car.cfl == 0.4 and car.r_hold_cell == 0.1
