"""Server policy on the CPU: admission by free device memory in arrival
order, retirement on the residual test or the step cap, the active-cell
cap, and the latency statistics.  The batch context is replaced by a model
whose clock advances with the work it is given, so these tests run without
a GPU (the GPU behaviour of the same loop is in test_gpu_serve.py)."""
import numpy as np
import pytest

from set_amd import mesh, serve, workloads


class SimClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


class ModelContext:
    """setmhd_batch_* model: a request of n cells converges after conv[n]
    steps; a queued step of the slots admitted at launch ends cells / rate
    seconds after it started on the shared clock; removed slots free their
    memory at the next wait, as the library does."""

    def __init__(self, clock, free_bytes, conv, rate=1e9):
        self.clock, self.free, self.conv, self.rate = clock, free_bytes, conv, rate
        self.slots = {}
        self.retired = {}
        self.log = []
        self.flight = None  # (end time, steps, slots)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False

    def mem_info(self):
        return self.free, 1 << 40

    def batch_add(self, grid, state):
        s = max(list(self.slots) + list(self.retired), default=-1) + 1
        self.slots[s] = [grid.n_cell, 0]
        self.free -= grid.n_cell * serve.BYTES_PER_CELL
        self.log.append(("add", grid.n_cell, self.clock()))
        return s

    def batch_remove(self, s):
        assert self.flight is None or s not in self.flight[2]
        self.retired[s] = self.slots.pop(s)
        self.log.append(("remove", self.retired[s][0], self.clock()))

    def batch_step_async(self, n):
        assert self.flight is None
        cells = sum(c for c, _ in self.slots.values())
        self.flight = (self.clock() + n * cells / self.rate, n, set(self.slots))

    def batch_poll(self):
        return self.flight is None or self.clock() >= self.flight[0]

    def batch_wait(self):
        if self.flight:
            end, n, ss = self.flight
            self.clock.t = max(self.clock.t, end)
            for s in ss:
                self.slots[s][1] += n
            self.flight = None
        for n, _ in self.retired.values():
            self.free += n * serve.BYTES_PER_CELL
        self.retired.clear()

    def batch_query(self, s):
        assert self.flight is None or s not in self.flight[2]
        n, steps = self.slots.get(s) or self.retired[s]
        return {"steps": steps, "t": 0.0, "dt": 1e-4,
                "residual": 1e-7 if steps >= self.conv[n] else 1e-3}

    def batch_download(self, s):
        return np.zeros(((self.slots.get(s) or self.retired[s])[0], 16)), 0.0

    # the copy of a result to the host lands copy_s seconds after it is queued
    copy_s = 0.0

    def batch_download_async(self, s):
        n = (self.slots.get(s) or self.retired[s])[0]
        self.tickets = getattr(self, "tickets", {})
        k = len(self.tickets)
        self.tickets[k] = (self.clock() + self.copy_s, n)
        self.log.append(("copy", n, self.clock()))
        return k

    def batch_download_view(self, k):
        ready, n = self.tickets[k]
        return None if self.clock() < ready else (np.zeros((n, 16)), 0.0)

    def batch_download_release(self, k):
        del self.tickets[k]

    def batch_download_collect(self, k, wait=True):
        ready, n = self.tickets[k]
        if self.clock() < ready:
            if not wait:
                return None
            self.clock.t = ready
        del self.tickets[k]
        return np.zeros((n, 16)), 0.0

    def batch_autotune(self, probe_steps=2):
        return (32, 8, 2)


def requests(sizes, arrivals):
    out = []
    for i, (n, t) in enumerate(zip(sizes, arrivals)):
        g = mesh.grid_2d(n, 1)
        out.append(workloads.Request(i, t, g, mesh.Inflow()))
    return out


def run(reqs, free, conv, **kw):
    clock = SimClock()
    ctx = ModelContext(clock, free, conv)
    srv = serve.Server(check_every=50, max_steps=kw.pop("max_steps", 5000), autotune=False,
                       reserve_bytes=0, context=lambda: ctx, **kw)
    states = [None] * len(reqs)
    done, stats = srv.run(reqs, states, clock=clock, sleep=clock.sleep)
    return done, stats, ctx


def test_admission_waits_for_free_memory():
    reqs = requests([1000, 2000, 3000], [0.0, 0.0, 0.0])
    conv = {1000: 100, 2000: 200, 3000: 300}
    # room for the first two together; the third fits only once both left
    done, stats, ctx = run(reqs, 3500 * serve.BYTES_PER_CELL, conv)
    adds = [e for e in ctx.log if e[0] == "add"]
    removes = [e for e in ctx.log if e[0] == "remove"]
    assert [a[1] for a in adds] == [1000, 2000, 3000]  # arrival order
    assert adds[0][2] == adds[1][2] == 0.0
    assert [r[1] for r in removes[:2]] == [1000, 2000]
    assert adds[2][2] >= removes[1][2] > removes[0][2]
    assert stats["solves"] == 3 and stats["converged"] == 3


def test_all_admitted_when_memory_allows_and_retired_on_convergence():
    reqs = requests([1000, 2000, 3000], [0.0, 0.0, 0.0])
    conv = {1000: 120, 2000: 260, 3000: 40}
    done, stats, ctx = run(reqs, 1 << 40, conv)
    by = {r["cells"]: r for r in done}
    for n, c in conv.items():
        # retired at the first check (every 50 steps) at or after convergence
        assert by[n]["steps"] == 50 * int(np.ceil(c / 50))
        assert by[n]["converged"]
    assert [r["cells"] for r in done] == [3000, 1000, 2000]


def test_step_cap_retires_unconverged():
    reqs = requests([1000], [0.0])
    done, stats, _ = run(reqs, 1 << 40, {1000: 10 ** 9}, max_steps=300)
    assert done[0]["steps"] == 300 and not done[0]["converged"]
    assert stats["converged"] == 0


def test_active_cell_cap_queues_in_arrival_order():
    reqs = requests([1000, 1000, 1000], [0.0, 0.0, 0.0])
    conv = {1000: 50}
    done, stats, ctx = run(reqs, 1 << 40, conv, max_active_cells=2000)
    live, peak = 0, 0
    for e in ctx.log:
        live += e[1] if e[0] == "add" else -e[1]
        peak = max(peak, live)
    assert peak == 2000 and stats["solves"] == 3


def test_arrival_during_a_chunk_is_admitted_before_it_ends():
    # 10^7-cell request: a 50-step chunk takes 0.5 s of model time; the second
    # request arrives 0.1 s in and is uploaded while that chunk runs
    reqs = requests([10 ** 7, 1000], [0.0, 0.1])
    done, stats, ctx = run(reqs, 1 << 40, {10 ** 7: 200, 1000: 100})
    adds = [e for e in ctx.log if e[0] == "add"]
    assert adds[1][2] < 0.5
    assert stats["solves"] == 2 and stats["converged"] == 2


def test_late_arrival_waits_and_latency_stats():
    reqs = requests([1000, 1000], [0.0, 5.0])
    done, stats, _ = run(reqs, 1 << 40, {1000: 100})
    lat = sorted(r["done"] - r["arrival"] for r in done)
    assert done[1]["admitted"] >= 5.0
    assert stats["latency_p50_s"] == pytest.approx(np.percentile(lat, 50))
    assert stats["latency_p99_s"] == pytest.approx(np.percentile(lat, 99))
    assert stats["solves_per_hour"] == pytest.approx(3600 * 2 / stats["wall_s"])


def mixed_run(split):
    """One large 3-D request and three small 2-D ones, all at t = 0; every
    lane gets its own model context on the shared clock."""
    big = workloads.Request(0, 0.0, mesh.grid_3d(100, 100, 100), mesh.Inflow())
    small = [workloads.Request(i, 0.0, mesh.grid_2d(1000, 1), mesh.Inflow()) for i in (1, 2, 3)]
    reqs = [big] + small
    conv = {10 ** 6: 1000, 1000: 100}
    clock = SimClock()
    made = []

    def factory():
        made.append(ModelContext(clock, 1 << 40, conv))
        return made[-1]

    srv = serve.Server(check_every=50, max_steps=5000, autotune=False, reserve_bytes=0,
                       context=factory, split_dims=split)
    done, stats = srv.run(reqs, [None] * 4, clock=clock, sleep=clock.sleep)
    return {r["rid"]: r for r in done}, stats, made


def test_split_lanes_let_small_requests_finish_first():
    one, _, made1 = mixed_run(False)
    two, _, made2 = mixed_run(True)
    assert len(made1) == 1 and len(made2) == 2
    # the same steps for every request either way
    assert {k: r["steps"] for k, r in one.items()} == {k: r["steps"] for k, r in two.items()}
    assert all(r["converged"] for r in two.values())
    # alone in their lane the 2-D requests no longer step at the 3-D grid's pace
    for rid in (1, 2, 3):
        assert two[rid]["done"] < 0.5 * one[rid]["done"]


def test_split_lanes_route_by_dimensionality():
    _, _, made = mixed_run(True)
    cells = [sorted(n for e, n, _ in ctx.log if e == "add") for ctx in made]
    assert sorted(cells) == [[1000, 1000, 1000], [10 ** 6]]


def test_retired_results_copy_beside_the_next_chunk():
    """A retired request's result is copied to the host while the lane's next
    chunk runs: the loop queues the copy, launches the chunk and collects the
    copy by polling, so the request's completion time is its copy's landing,
    not the end of the chunk, and nothing waits on the copy."""
    reqs = requests([1000, 400000], [0.0, 0.0])
    clock = SimClock()
    ctx = ModelContext(clock, 1 << 40, {1000: 50, 400000: 10 ** 9})
    ctx.copy_s = 1e-3
    srv = serve.Server(check_every=50, max_steps=200, autotune=False, reserve_bytes=0,
                       context=lambda: ctx, split_dims=False)
    done, stats = srv.run(reqs, [None, None], clock=clock, sleep=clock.sleep)
    small = next(d for d in done if d["cells"] == 1000)
    t_copy = next(t for kind, n, t in ctx.log if kind == "copy" and n == 1000)
    chunk = 50 * 400000 / ctx.rate  # the big request's next chunk alone
    assert small["done"] >= t_copy + ctx.copy_s
    assert small["done"] < t_copy + chunk  # collected before that chunk ended
    assert stats["solves"] == 2
