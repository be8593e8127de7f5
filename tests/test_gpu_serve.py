"""Continuous-batching server: every request's result is the same whether it
ran alone (one-at-a-time path) or batched with whatever else was on the GPU;
converged requests leave early; admission respects the memory reserve."""
import numpy as np
import pytest

from conftest import rel_err
from set_amd import mesh, serve, workloads

pytestmark = pytest.mark.gpu


def small_requests(n=7, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(n):
        t += float(rng.exponential(0.01))
        if i % 3 == 2:
            m = int(rng.integers(20, 33))
            g = mesh.grid_3d(m, m, m)
        else:
            s = float(rng.uniform(1.0, 2.0))
            g = mesh.grid_2d(int(128 * s), int(64 * s))
        out.append(workloads.Request(i, t, g, workloads._params(rng)))
    return out


@pytest.mark.parametrize("split_dims", [True, False])
def test_each_result_same_alone_and_in_the_mix(split_dims):
    print(f"[case] serving: batched (split_dims={split_dims}) vs one-at-a-time results", flush=True)
    reqs = small_requests()
    states = [r.state() for r in reqs]
    kw = dict(check_every=20, max_steps=60, tol=1e-300)
    done_b, stats_b = serve.Server(autotune=True, split_dims=split_dims, **kw).run(
        reqs, states, keep_results=True)
    done_s, stats_s = serve.run_sequential(reqs, states, keep_results=True, **kw)
    assert stats_b["solves"] == stats_s["solves"] == len(reqs)
    by_b = {d["rid"]: d for d in done_b}
    for d in done_s:
        b = by_b[d["rid"]]
        assert b["steps"] == d["steps"] == 60
        e = rel_err(mesh.prim_to_cons(b["result"][:, :8]), mesh.prim_to_cons(d["result"][:, :8]), "inf")
        assert np.all(e <= 1e-12), (d["rid"], e)
        assert b["t"] == pytest.approx(d["t"], rel=1e-13)
    print("batched", stats_b, "sequential", stats_s)


def test_converged_request_leaves_early():
    print("[case] serving: a steady state converges at the first check", flush=True)
    g = mesh.grid_2d(128, 64)
    W = np.zeros((g.n_cell, 16))
    W[:, :8] = [0.1, -2.0, 0.0, 0.0, 0.1, 0.2, 0.1, 0.0]  # uniform: residual exactly 0
    req = [workloads.Request(0, 0.0, g, mesh.Inflow())]
    srv = serve.Server(check_every=10, max_steps=5000, tol=1e-6, autotune=False)
    done, stats = srv.run(req, [W])
    assert done[0]["converged"] and done[0]["steps"] == 10 and done[0]["residual"] == 0.0
