"""Continuous-batching server for heliosphere solve requests.

Requests arrive over time (a seeded Poisson stream); each is a structured
grid with its own inflow and field.  The server keeps every admitted
request's state resident on the GPU as one slot of a batch context
(setmhd_batch_*): all active requests advance together, one persistent
fused launch per step per dimensionality, replayed from a captured graph.
Around chunks of steps it

  * admits arrived requests in arrival order while free HBM (hipMemGetInfo)
    covers their state plus a reserve (uploads overlap the running chunk),
  * checks each request's relative density residual
    ||rho^{n+1} - rho^n||_1 / ||rho^n||_1 of its last step, and
  * retires a request when the residual is <= tol or it reached max_steps,
    downloading its state as the result while the next chunk runs.

A request that arrives while others are on the GPU therefore starts at the
next chunk boundary instead of waiting for the running ones to finish.
run_sequential is the one-at-a-time baseline: each request is uploaded,
solved alone with the single-grid path and downloaded before the next.
"""
import collections
import time

import numpy as np

from . import solver

# device bytes per cell of an admitted grid: two FP64 state buffers,
# Cell_par(9:16), and room for the library's Cell_par staging (16 doubles)
BYTES_PER_CELL = 64 + 64 + 64 + 128


def _latency_stats(done, wall):
    lat = np.array([r["done"] - r["arrival"] for r in done])
    return {
        "solves": len(done),
        "wall_s": wall,
        "solves_per_hour": 3600.0 * len(done) / wall if wall > 0 else 0.0,
        "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "converged": int(sum(r["converged"] for r in done)),
        "cell_updates": float(sum(r["cells"] * r["steps"] for r in done)),
    }


class Server:
    def __init__(self, precision=solver.PREC_FP64, check_every=50, max_steps=5000, tol=1e-6,
                 reserve_bytes=4 << 30, max_active=512, max_active_cells=None, autotune=False,
                 context=None, poll_s=0.001):
        self.precision = precision
        self.check_every = check_every
        self.max_steps = max_steps
        self.tol = tol
        self.reserve = reserve_bytes
        self.max_active = max_active
        # optional cap on the cells stepping at once (beyond what fills the
        # GPU, more concurrency only stretches every request's latency);
        # a request larger than the cap still runs, alone
        self.max_active_cells = max_active_cells
        # Off by default: the library's launches (2-D: persistent blocks;
        # 3-D: one block per tile) measured better on a changing request mix
        # than a tuning done on the first two requests.
        self.autotune = autotune
        self.poll_s = poll_s  # host sleep while a chunk runs and requests are due
        # context factory (a solver.Context by default; tests pass a model)
        self.context = context or (lambda: solver.Context(solver.cartesian_config(precision=precision)))

    def run(self, requests, states, clock=time.perf_counter, sleep=time.sleep, keep_results=False):
        """requests: workloads.Request list; states: their initial Cell_par
        (same order).  Returns per-request records and summary stats.

        Each pass of the loop admits arrived requests (their uploads run on
        the library's side stream, beside the chunk in flight), waits for
        the chunk, retires converged or capped requests, starts the next
        chunk, and downloads the retired results while it runs."""
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        pending = collections.deque(order)
        active = {}  # slot -> [request index, admitted time, steps done]
        done = []
        tuned = not self.autotune
        spent = collections.Counter()  # seconds per server activity
        active_cells = 0
        with self.context() as ctx:
            t0 = clock()
            in_flight = False
            while pending or active or in_flight:
                now = clock() - t0
                free, _ = ctx.mem_info()
                while (pending and requests[pending[0]].arrival <= now
                       and len(active) < self.max_active):
                    i = pending[0]
                    need = requests[i].cells * BYTES_PER_CELL
                    if free - need < self.reserve:
                        break  # admitted when memory frees up
                    if (self.max_active_cells is not None and active
                            and active_cells + requests[i].cells > self.max_active_cells):
                        break  # admitted when running requests retire
                    pending.popleft()
                    ta = clock()
                    slot = ctx.batch_add(requests[i].grid, states[i])
                    spent["admit"] += clock() - ta
                    active[slot] = [i, clock() - t0, 0]
                    active_cells += requests[i].cells
                    free -= need
                retiring = []
                if in_flight and pending and not ctx.batch_poll():
                    # keep admitting arrivals while the chunk runs
                    nxt = requests[pending[0]].arrival - (clock() - t0)
                    sleep(min(nxt, self.poll_s) if nxt > 0.0 else self.poll_s)
                    continue
                if in_flight:
                    ta = clock()
                    ctx.batch_wait()
                    spent["step"] += clock() - ta
                    in_flight = False
                    for slot in list(active):
                        q = ctx.batch_query(slot)
                        active[slot][2] = q["steps"]
                        if q["steps"] == 0:
                            continue  # admitted while the chunk ran
                        conv = q["residual"] <= self.tol
                        if conv or q["steps"] >= self.max_steps:
                            ctx.batch_remove(slot)  # readable until the next wait
                            i, admitted, steps = active.pop(slot)
                            active_cells -= requests[i].cells
                            retiring.append((slot, i, admitted, steps, q["residual"], conv))
                if active:
                    if not tuned and len(active) > 1:
                        ta = clock()
                        ctx.batch_autotune(probe_steps=2)
                        spent["autotune"] += clock() - ta
                        tuned = True
                    chunk = min(self.check_every,
                                min(self.max_steps - a[2] for a in active.values()))
                    ctx.batch_step_async(chunk)
                    in_flight = True
                for slot, i, admitted, steps, res, conv in retiring:
                    ta = clock()
                    cp, t = ctx.batch_download(slot)  # overlaps the chunk just started
                    spent["retire"] += clock() - ta
                    done.append({"rid": requests[i].rid, "arrival": requests[i].arrival,
                                 "admitted": admitted, "done": clock() - t0, "steps": steps,
                                 "residual": res, "converged": conv, "cells": requests[i].cells,
                                 "t": t, "result": cp if keep_results else None})
                if not active and not in_flight and pending:
                    wait = requests[pending[0]].arrival - (clock() - t0)
                    if wait > 0.0:
                        sleep(wait)
                    elif retiring:
                        ctx.batch_wait()  # frees the retired requests' memory; admit again
                    else:  # arrived, nothing running, still not admitted
                        raise RuntimeError(f"request {requests[pending[0]].rid} does not fit in "
                                           "device memory")
            wall = clock() - t0
        stats = _latency_stats(done, wall)
        stats["seconds_in"] = dict(spent)
        return done, stats


def run_sequential(requests, states, precision=solver.PREC_FP64, check_every=50, max_steps=5000,
                   tol=1e-6, clock=time.perf_counter, sleep=time.sleep, keep_results=False):
    """One request at a time, in arrival order, with the single-grid path."""
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    done = []
    t0 = clock()
    for i in order:
        r = requests[i]
        wait = r.arrival - (clock() - t0)
        if wait > 0:
            sleep(wait)
        admitted = clock() - t0
        with solver.Context(solver.cartesian_config(precision=precision)) as c:
            c.upload_structured(r.grid, states[i])
            c.dt_from_state()
            steps, res = 0, 1.0
            while True:
                n = min(check_every, max_steps - steps)
                if n > 1:
                    c.step(n - 1)
                res = c.step_residual()
                steps += n
                if res <= tol or steps >= max_steps:
                    break
            cp, t = c.download()
        done.append({"rid": r.rid, "arrival": r.arrival, "admitted": admitted,
                     "done": clock() - t0, "steps": steps, "residual": res,
                     "converged": res <= tol, "cells": r.cells, "t": t,
                     "result": cp if keep_results else None})
    wall = clock() - t0
    return done, _latency_stats(done, wall)
