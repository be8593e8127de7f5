"""Continuous-batching server for heliosphere solve requests.

Requests arrive over time (a seeded Poisson stream); each is a structured
grid with its own inflow and field.  The server keeps every admitted
request's state resident on the GPU as one slot of a batch context
(setmhd_batch_*): all active requests of a lane advance together, one
fused launch per step, replayed from a captured graph.  2-D and 3-D
requests run in two lanes (two contexts, two streams).  Around chunks of
steps it

  * admits arrived requests in arrival order while free HBM (hipMemGetInfo)
    covers their state plus a reserve (uploads overlap the running chunk),
  * checks each request's relative density residual
    ||rho^{n+1} - rho^n||_1 / ||rho^n||_1 of its last step, and
  * retires a request when the residual is <= tol or it reached max_steps:
    its state is copied to pinned host memory on the library's second stream
    while the next chunk runs, and collected by polling (the loop never
    blocks on a result).

A request that arrives while others are on the GPU therefore starts at the
next chunk boundary instead of waiting for the running ones to finish.
run_sequential is the one-at-a-time baseline: each request is uploaded,
solved alone with the single-grid path and downloaded before the next.
"""
import collections
import contextlib
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


class _Lane:
    """One batch context and the requests routed to it."""

    def __init__(self, ctx, pending):
        self.ctx = ctx
        self.pending = collections.deque(pending)  # request indices, arrival order
        self.active = {}  # slot -> [request index, admitted time, steps done]
        self.in_flight = False
        self.active_cells = 0
        self.tuned = False
        self.downloads = []  # [ticket, record] of results being copied to the host
        self.retired_held = False  # retired requests' memory not yet freed

    def busy(self):
        return bool(self.pending or self.active or self.in_flight or self.downloads)


class Server:
    def __init__(self, precision=solver.PREC_FP64, check_every=50, max_steps=5000, tol=1e-6,
                 reserve_bytes=4 << 30, max_active=512, max_active_cells=None, autotune=False,
                 context=None, poll_s=0.001, split_dims=True):
        self.precision = precision
        self.check_every = check_every
        self.max_steps = max_steps
        self.tol = tol
        self.reserve = reserve_bytes
        self.max_active = max_active
        # optional cap on the cells stepping at once per lane (beyond what
        # fills the GPU, more concurrency only stretches every request's
        # latency); a request larger than the cap still runs, alone
        self.max_active_cells = max_active_cells
        # Off by default: the library's launches (one block per tile)
        # measured better on a changing request mix than a tuning done on
        # the first two requests.
        self.autotune = autotune
        self.poll_s = poll_s  # host sleep while chunks run
        # 2-D and 3-D requests in two batch contexts (two streams): the
        # small 2-D grids advance at their own step rate instead of the
        # large 3-D grids' (one context when False)
        self.split_dims = split_dims
        # context factory (a solver.Context by default; tests pass a model)
        self.context = context or (lambda: solver.Context(solver.cartesian_config(precision=precision)))

    def _admit(self, lane, requests, states, now, free, clock, t0, spent):
        """Admits lane's arrived requests in arrival order while memory and
        the caps allow; returns the free bytes left."""
        while (lane.pending and requests[lane.pending[0]].arrival <= now
               and len(lane.active) < self.max_active):
            i = lane.pending[0]
            need = requests[i].cells * BYTES_PER_CELL
            if free - need < self.reserve:
                break  # admitted when memory frees up
            if (self.max_active_cells is not None and lane.active
                    and lane.active_cells + requests[i].cells > self.max_active_cells):
                break  # admitted when running requests retire
            lane.pending.popleft()
            ta = clock()
            slot = lane.ctx.batch_add(requests[i].grid, states[i])
            spent["admit"] += clock() - ta
            lane.active[slot] = [i, clock() - t0, 0]
            lane.active_cells += requests[i].cells
            free -= need
        return free

    def _complete(self, lane, requests, spent, clock):
        """Completes lane's chunk and removes its converged or capped
        requests; returns them (readable until the lane's next wait)."""
        ta = clock()
        lane.ctx.batch_wait()
        spent["step"] += clock() - ta
        lane.in_flight = False
        retiring = []
        for slot in list(lane.active):
            q = lane.ctx.batch_query(slot)
            lane.active[slot][2] = q["steps"]
            if q["steps"] == 0:
                continue  # admitted while the chunk ran
            conv = q["residual"] <= self.tol
            if conv or q["steps"] >= self.max_steps:
                lane.ctx.batch_remove(slot)
                i, admitted, steps = lane.active.pop(slot)
                lane.active_cells -= requests[i].cells
                retiring.append((slot, i, admitted, steps, q["residual"], conv))
        return retiring

    def _collect(self, lane, done, clock, t0, spent, keep):
        """Collects the results whose copies have landed (never blocks)."""
        got = False
        for item in list(lane.downloads):
            ticket, rec = item
            ta = clock()
            r = lane.ctx.batch_download_view(ticket)  # the landed result, read in place
            if r is None:
                spent["retire_collect"] += clock() - ta
                continue
            view, t = r
            cp = np.array(view) if keep else None
            lane.ctx.batch_download_release(ticket)
            spent["retire_collect"] += clock() - ta
            lane.downloads.remove(item)
            done.append(dict(rec, done=clock() - t0, t=t, result=cp))
            got = True
        return got

    def _launch(self, lane, spent, clock):
        if not lane.tuned and len(lane.active) > 1:
            ta = clock()
            lane.ctx.batch_autotune(probe_steps=2)
            spent["autotune"] += clock() - ta
            lane.tuned = True
        chunk = min(self.check_every, min(self.max_steps - a[2] for a in lane.active.values()))
        lane.ctx.batch_step_async(chunk)
        lane.in_flight = True

    def run(self, requests, states, clock=time.perf_counter, sleep=time.sleep, keep_results=False):
        """requests: workloads.Request list; states: their initial Cell_par
        (same order).  Returns per-request records and summary stats.

        Each pass of the loop, per lane: admits arrived requests (their
        uploads run on the library's side stream, beside the chunk in
        flight); when the lane's chunk has ended, retires converged or
        capped requests, starts the next chunk, and downloads the retired
        results while it runs."""
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        keys = [0, 1] if self.split_dims else [0]
        route = (lambda i: int(requests[i].grid.nz > 1)) if self.split_dims else (lambda i: 0)
        done = []
        spent = collections.Counter()  # seconds per server activity
        with contextlib.ExitStack() as stack:
            lanes = []
            for k in keys:
                mine = [i for i in order if route(i) == k]
                if mine:
                    lanes.append(_Lane(stack.enter_context(self.context()), mine))
            for lane in lanes:
                lane.tuned = not self.autotune
            t0 = clock()
            while any(lane.busy() for lane in lanes):
                free, _ = lanes[0].ctx.mem_info()
                progressed = False
                for lane in lanes:
                    free = self._admit(lane, requests, states, clock() - t0, free, clock, t0, spent)
                    if self._collect(lane, done, clock, t0, spent, keep_results):
                        progressed = True
                    retiring = []
                    if lane.in_flight:
                        # block only when nothing else can need the host
                        alone = (all(not o.busy() for o in lanes if o is not lane) and not lane.pending
                                 and not lane.downloads)
                        if not alone and not lane.ctx.batch_poll():
                            continue
                        retiring = self._complete(lane, requests, spent, clock)
                        progressed = True
                    if lane.active and not lane.in_flight:
                        self._launch(lane, spent, clock)
                        progressed = True
                    for slot, i, admitted, steps, res, conv in retiring:
                        ta = clock()
                        ticket = lane.ctx.batch_download_async(slot)  # beside the chunk just started
                        spent["retire_queue"] += clock() - ta
                        lane.downloads.append([ticket, {
                            "rid": requests[i].rid, "arrival": requests[i].arrival,
                            "admitted": admitted, "steps": steps, "residual": res, "converged": conv,
                            "cells": requests[i].cells}])
                    if not lane.active and not lane.in_flight and not lane.downloads and lane.retired_held:
                        lane.ctx.batch_wait()  # frees the retired requests' memory
                        lane.retired_held = False
                    if retiring:
                        lane.retired_held = True
                if progressed:
                    continue
                now = clock() - t0
                arrivals = [requests[lane.pending[0]].arrival for lane in lanes if lane.pending]
                nxt = min(arrivals) - now if arrivals else None
                if any(lane.in_flight or lane.downloads for lane in lanes):
                    sleep(min(nxt, self.poll_s) if nxt is not None and nxt > 0.0 else self.poll_s)
                elif nxt is not None and nxt > 0.0:
                    sleep(nxt)
                else:  # arrived, nothing running, still not admitted
                    lane = next(lane for lane in lanes if lane.pending)
                    raise RuntimeError(f"request {requests[lane.pending[0]].rid} does not fit in "
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
