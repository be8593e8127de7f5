"""bench.py's host side, without a GPU: the multi-GPU refusal, the figures
it reads from the committed profiles, and its error table."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_profile_figures_are_committed():
    """The roofline's traffic, the flops and the VALU-issue ceiling come from
    the committed rocprofv3 summary of the headline kernel."""
    t = bench.from_profiles("fused_fp64_euler_256cube", "traffic_bytes_per_launch")
    f = bench.from_profiles("fused_fp64_euler_256cube", "fp64_flops_per_cell_update")
    v = bench.from_profiles("fused_fp64_euler_256cube", "valu_instructions_per_cell_update")
    alg = 256 ** 3 * 128
    assert alg <= t <= 1.5 * alg
    assert 500 < f < 2000 and 500 < v < 2000
    assert bench.from_profiles("fused_fp64_euler_256cube", "no_such_key") is None
    assert bench.from_profiles("no_such_profile", "traffic_bytes_per_launch") is None


def test_error_table():
    b = np.ones((10, 8))
    a = b.copy()
    a[3, 0] += 0.5  # rho: L1 0.5 / 10, L-inf 0.5 / 1
    a[:, 7] *= 1.0 + 1e-9
    e = bench.error_table(a, b)
    assert e["rel_l1"]["rho"] == 0.05 and e["rel_linf"]["rho"] == 0.5
    assert abs(e["rel_l1"]["bz"] - 1e-9) < 1e-15 and e["rel_l1"]["mx"] == 0.0
