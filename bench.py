"""Throughput of the fused ideal-MHD step on one MI355X.

Workload: the heliosphere set-up on a
3-D 256^3 box [-1,1]^3 -- supersonic radial wind held inside r < 0.1, a
uniform magnetised inflow (rho 0.1 with +-1 % seeded noise, u = -2.5,
p = 0.1, |B| = 0.3 tilted 30 deg in x-y) elsewhere -- first-order HLLD,
forward Euler, CFL 0.4, outflow boundaries, FP64.  Synthetic data.

A step is one fused flux + update launch over all 16.8 M cells.  `value` is
cell-updates/s over the timed steps (wall clock between device syncs, state
resident in HBM).  The roofline object uses the kernel's own duration from
HIP events on the library's stream and 128 algorithmic bytes per
cell-update (read U^n, write U^{n+1}: 8 FP64 each).  The mixed-precision
and Heun-RK2 forms, the 64-grid batch, the serving stream, the drop-in
face-list path at reference semantics, and the face-list path on this same
box are reported beside it under "forms".  cpu_baseline is
the CPU oracle (the C restatement of the reference algorithm) on a bounded
sample of the same workload.

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--n 256]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "HLLD cell-updates/s FP64 vs mixed (% HBM peak); p50/p99 solve latency"


PEAK_FP64_TFLOPS = 78.6


def from_profiles(name, key):
    """A figure measured by rocprofv3 PMC runs, when a summary for this
    kernel/workload is committed (profiles/<name>.txt holds 'key: value'),
    e.g. traffic_bytes_per_launch or fp64_flops_per_cell_update."""
    p = os.path.join(ROOT, "profiles", name + ".txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith(key + ":"):
            return int(float(line.split(":")[1]))
    return None


def run_form(g, W, precision, time_scheme, steps, warmup, spin_steps=0):
    """One form on grid g from Cell_par W: `warmup` untimed steps, then
    `spin_steps` more untimed steps (a device spin-up, so that the timed
    steps do not run while the clocks ramp; the same count for every form),
    then exactly `steps` timed steps.  The timed region is bracketed by host
    clocks and device syncs (value) and, on the library's stream, by HIP
    events around the same launches (the kernel's average launch duration)."""
    from set_amd import solver
    cfg = solver.cartesian_config(cfl=0.4, r_hold=0.1, precision=precision,
                                  time_scheme=time_scheme)
    with solver.Context(cfg) as c:
        t0 = time.perf_counter()
        c.upload_structured(g, W)
        c.sync()
        t_up = time.perf_counter() - t0
        c.dt_from_state()
        t0 = time.perf_counter()
        if warmup:
            c.step(warmup)
        c.sync()
        t_warm = time.perf_counter() - t0
        if spin_steps:
            c.step(spin_steps)
            c.sync()
        t0 = time.perf_counter()
        ms_ev = c.step_timed(steps)  # HIP events on the library's stream, these launches
        c.sync()
        wall = time.perf_counter() - t0
        t_sim, dt = c.sync()
        t0 = time.perf_counter()
        cp, _ = c.download()
        t_down = time.perf_counter() - t0
    del cp
    launches = steps * (2 if time_scheme == 1 else 1)
    bytes_per_cell = 320 if time_scheme == 1 else 128
    return {
        "value": g.n_cell * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "kernel_ms_per_launch": ms_ev / launches,
        "achieved_GBs": g.n_cell * bytes_per_cell * steps / (ms_ev * 1e-3) / 1e9,
        "achieved_GBs_wall": g.n_cell * bytes_per_cell * steps / wall / 1e9,
        # the same steps with one upload and one download of Cell_par
        # (host buffers over PCIe, conversion kernels included)
        "upload_s": t_up,
        "download_s": t_down,
        "pcie_inclusive_value": g.n_cell * steps / (wall + t_up + t_down),
        "t_sim": t_sim,
        "dt": dt,
        "warmup_s": t_warm,
        "steps_run": warmup + spin_steps + steps,
    }


def precision_error(g, W, steps):
    """The mixed form against FP64 after `steps` steps from the same initial
    state (untimed runs): per-variable relative L1 and L-inf errors."""
    from set_amd import mesh, solver
    out = []
    for precision in (0, 1):
        with solver.Context(solver.cartesian_config(cfl=0.4, r_hold=0.1, precision=precision)) as c:
            c.upload_structured(g, W)
            c.dt_from_state()
            c.step(steps)
            cp, _ = c.download()
        out.append(mesh.prim_to_cons(cp[:, :8]))
        del cp
    return {"steps": steps, **error_table(out[1], out[0])}


def error_table(a, b):
    """Per-variable normwise relative L1 and L-inf errors of conserved state a
    against b: sum|a - b| / sum|b| and max|a - b| / max|b| per variable."""
    import numpy as np
    names = ("rho", "mx", "my", "mz", "E", "bx", "by", "bz")
    d = np.abs(a - b)
    l1 = d.sum(0) / np.maximum(np.abs(b).sum(0), 1e-300)
    li = d.max(0) / np.maximum(np.abs(b).max(0), 1e-300)
    return {"rel_l1": {n: float(v) for n, v in zip(names, l1)},
            "rel_linf": {n: float(v) for n, v in zip(names, li)}}


def batch_form(precision, steps, seq_steps):
    """64 uneven 2-D grids in one persistent launch per step (graph-replayed,
    autotuned tiles) against the same grids stepped one at a time."""
    from set_amd import solver, workloads
    grids = workloads.batch_grids(64)
    cells = sum(g.n_cell for g, _ in grids)
    with solver.Context(solver.cartesian_config(precision=precision)) as c:
        slots = [c.batch_add(g, W) for g, W in grids]
        tile = c.batch_autotune(probe_steps=5)
        c.batch_step(10)
        c.batch_query(slots[0])
        t0 = time.perf_counter()
        c.batch_step(steps)
        c.batch_query(slots[0])
        wall = time.perf_counter() - t0
    seq = 0.0
    for g, W in grids:
        with solver.Context(solver.cartesian_config(precision=precision)) as c:
            c.upload_structured(g, W)
            c.dt_from_state()
            c.step(3)
            c.sync()
            t0 = time.perf_counter()
            c.step(seq_steps)
            c.sync()
            seq += (time.perf_counter() - t0) / seq_steps
    value = cells * steps / wall
    return {"value": value, "grids": len(grids), "cells": cells, "steps": steps,
            "ms_per_step": 1e3 * wall / steps,
            "launch_2d": list(tile[2]),  # tile_x, tile_y, blocks per CU (0: one per tile)
            "hbm_frac": value * 128 / (PEAK_HBM_GBS * 1e9),
            "sequential_value": cells / seq, "speedup_vs_one_at_a_time": (cells / seq and value / (cells / seq))}


def serving_form(n_requests, max_steps, precision=0, burst=True):
    """The serving stream through the continuous-batching server and one at a
    time.  work_bound: both modes' cell-update rates over the same work.
    The one-at-a-time rate is over its busy time (the sum of its solves'
    durations), so it is what it reaches however dense the stream; while
    both modes are bound by the work rather than the arrivals, the
    solves/hour ratio is the ratio of the two rates.  burst: the same
    requests all arriving at once, the server's capacity."""
    from set_amd import serve, workloads
    reqs = workloads.serving_requests(n_requests)
    states = [r.state() for r in reqs]
    _, b = serve.Server(precision=precision, max_steps=max_steps).run(reqs, states)
    done_q, q = serve.run_sequential(reqs, states, precision=precision, max_steps=max_steps)
    busy_q = sum(d["done"] - d["admitted"] for d in done_q)
    work = b["cell_updates"]
    work_3d = sum(d["cells"] * d["steps"] for d, r in zip(sorted(done_q, key=lambda d: d["rid"]),
                                                            sorted(reqs, key=lambda r: r.rid))
                  if r.grid.nz > 1)
    out = {"requests": n_requests, "max_steps": max_steps, "batched": b, "one_at_a_time": q,
           "solves_per_hour_ratio": b["solves_per_hour"] / q["solves_per_hour"],
           "work_bound": {"cell_updates": work, "frac_3d": work_3d / work if work else None,
                          "last_arrival_s": max(r.arrival for r in reqs),
                          "batched_rate": work / b["wall_s"],
                          "one_at_a_time_busy_rate": q["cell_updates"] / busy_q,
                          "rate_ratio": (work / b["wall_s"]) / (q["cell_updates"] / busy_q)}}
    if burst:
        # the same requests all arriving at once: the server's capacity, which
        # the stream's arrivals do not bound; one at a time is busy all along
        # either way, so its busy rate is its capacity
        import dataclasses
        now = [dataclasses.replace(r, arrival=0.0) for r in reqs]
        _, bb = serve.Server(precision=precision, max_steps=max_steps).run(now, states)
        out["burst"] = {"batched": bb, "batched_rate": bb["cell_updates"] / bb["wall_s"],
                        "capacity_ratio": (bb["cell_updates"] / bb["wall_s"]) / (q["cell_updates"] / busy_q)}
    return out


def small_form(nx=400, ny=200, steps=500):
    """The CPU-runnable 2-D case (400 x 200, the reference's own size class):
    the structured step in FP64 and mixed and the drop-in face list (HLLD,
    forward Euler), timed with HIP events over `steps` steps.  A step of so
    small a grid is bound by the latency of its row march, not by HBM."""
    from set_amd import mesh, solver
    g = mesh.grid_2d(nx, ny)
    W = mesh.heliosphere_state(g)
    out = {"grid": f"{nx}x{ny}", "cells": g.n_cell, "steps": steps}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        with solver.Context(solver.cartesian_config(precision=prec)) as c:
            c.upload_structured(g, W)
            c.dt_from_state()
            c.step(50)
            c.sync()
            ms = c.step_timed(steps)
        out[name] = {"us_per_step": ms * 1e3 / steps, "value": g.n_cell * steps / (ms * 1e-3)}
    with solver.Context(solver.cartesian_config()) as c:
        c.upload_mesh(mesh.face_mesh(g), W)
        c.dt_from_state()
        c.step(50)
        c.sync()
        ms = c.step_timed(steps)
    out["dropin_exact"] = {"us_per_step": ms * 1e3 / steps, "value": g.n_cell * steps / (ms * 1e-3)}
    return out


def dropin_form(nx, ny, steps):
    """The drop-in face-list path at SET's production semantics (HLLC with the
    LAX wedge, minmod TVD, Heun, axisymmetric and charge-exchange sources,
    w = B = 0, reference initial dt 2e-9) on a half-plane heliosphere mesh:
    the exact build (bit for bit with the oracle) and the fast build, each
    with two launches per Heun stage and with one (SETMHD_F_PER_CELL)."""
    import numpy as np

    from set_amd import mesh, solver, workloads
    fm, W = workloads.heliosphere_face_mesh(nx, ny)
    # the same mesh numbered at random (a mesh file in no useful order), run
    # as given and with SETMHD_F_REORDER (device numbering by Morton order)
    rng = np.random.default_rng(1)
    cp, gp = rng.permutation(fm.n_cell), rng.permutation(fm.n_gran)
    fr = mesh.renumber_face_mesh(fm, cp, gp)
    out = {"mesh": f"{nx}x{ny} half-plane heliosphere face list", "cells": fm.n_cell,
           "faces": fm.n_gran, "steps": steps, "time_scheme": "heun"}
    for name, m, w, flags in (("exact", fm, W, 0), ("fast", fm, W, solver.F_FAST),
                              ("exact_per_cell", fm, W, solver.F_PER_CELL),
                              ("fast_per_cell", fm, W, solver.F_FAST | solver.F_PER_CELL),
                              ("exact_random_numbering", fr, W[cp], 0),
                              ("exact_random_numbering_reorder", fr, W[cp], solver.F_REORDER)):
        cfg = solver.reference_config()
        cfg.flags |= flags
        with solver.Context(cfg) as c:
            c.upload_mesh(m, w)
            c.set_dt(2.0e-9)
            c.step(3)
            c.sync()
            ms = c.step_timed(steps) / steps
        out[name] = {"value": fm.n_cell / (ms * 1e-3), "unit": "cell-updates/s", "ms_per_step": ms}
    return out


def facelist_form(n, steps):
    """The drop-in face-list kernels (the reference's face pass and cell pass
    restated, with chlld's HLLD) on the headline workload itself: the n^3
    heliosphere box expressed as SET's face-list arrays, forward Euler, the
    same dt rule and held wind cells; exact build (bit for bit with the
    oracle) and fast build.  Beside the fused structured kernel this shows
    what the Cartesian specialisation buys on the same cells."""
    from set_amd import mesh, solver
    g = mesh.grid_3d(n, n, n)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g)
    out = {"mesh": f"{n}^3 box [-1,1]^3 as a face list", "cells": fm.n_cell, "faces": fm.n_gran,
           "steps": steps, "time_scheme": "euler"}
    for name, flags in (("exact", 0), ("fast", solver.F_FAST)):
        cfg = solver.cartesian_config(cfl=0.4, r_hold=0.1)
        cfg.flags |= flags
        with solver.Context(cfg) as c:
            c.upload_mesh(fm, W)
            c.dt_from_state()
            c.step(2)
            c.sync()
            ms = c.step_timed(steps) / steps
        out[name] = {"value": fm.n_cell / (ms * 1e-3), "unit": "cell-updates/s", "ms_per_step": ms}
    return out


def ena_form(nx, ny, n_v, cpu=True):
    """ENA line-of-sight spectra (SET_ENA.cpp:544-780, Culc_ENA): the
    reference's 180 angles (0.008 to pi - 0.008 rad by 1 degree) and n_v
    speeds through the triangulated cell centres of an nx x ny half-plane
    heliosphere.  Unit of work: one sample of one line of sight for one
    speed (the survival and intensity update of SET_ENA.cpp:875-921).  The
    CPU figure is the oracle (one thread, its brute-force triangle search
    included) on one of the angles."""
    import numpy as np

    import oracle
    from set_amd import ena, workloads
    fm, W = workloads.heliosphere_face_mesh(nx, ny)
    t0 = time.perf_counter()
    tri = ena.triangulate(fm.cell_center)
    t_tri = time.perf_counter() - t0
    angles = ena.map_angles()
    ena.spectra(fm.cell_center, W, angles[:4], 64, tri=tri)  # warm-up
    t0 = time.perf_counter()
    j, n = ena.spectra(fm.cell_center, W, angles, n_v, tri=tri)
    wall = time.perf_counter() - t0
    work = float(n.sum()) * n_v
    out = {"mesh": f"{nx}x{ny} half-plane heliosphere", "cells": fm.n_cell, "triangles": len(tri[0]),
           "angles": len(angles), "speeds": n_v, "samples": int(n.sum()), "seconds": wall,
           "value": work / wall, "unit": "sample-speed updates/s", "triangulation_s": t_tri}
    if cpu:
        k = len(angles) // 2
        t0 = time.perf_counter()
        ref, nref = oracle.ena_los(fm.cell_center, W, tri[0], angles[k], n_v)
        tc = time.perf_counter() - t0
        out["cpu_oracle_one_thread"] = {"value": nref * n_v / tc, "angle_rad": float(angles[k]), "seconds": tc}
        out["max_rel_diff_vs_oracle"] = float(np.abs(j[k] - ref).max() / max(ref.max(), 1e-300))
    return out


def copy_ceiling(nbytes, reps=10):
    """The device's plain copy rate on this box: a device-to-device
    hipMemcpy of nbytes (read nbytes + write nbytes, the traffic of one
    step) through the HIP runtime libsetmhd.so runs on, timed with HIP events
    after a warm-up.  A practical ceiling beside the nominal HBM peak."""
    import ctypes
    from set_amd import solver
    solver.load()
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime already loaded by libsetmhd.so
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]
    d2d = 3  # hipMemcpyDeviceToDevice

    def ok(st):
        if st != 0:
            raise RuntimeError(f"HIP error {st}")
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    ok(hip.hipMalloc(ctypes.byref(a), nbytes))
    try:
        ok(hip.hipMalloc(ctypes.byref(b), nbytes))
        try:
            ok(hip.hipEventCreate(ctypes.byref(e0)))
            ok(hip.hipEventCreate(ctypes.byref(e1)))
            for _ in range(3):
                ok(hip.hipMemcpy(b, a, nbytes, d2d))
            ok(hip.hipEventRecord(e0, None))
            for _ in range(reps):
                ok(hip.hipMemcpy(b, a, nbytes, d2d))
            ok(hip.hipEventRecord(e1, None))
            ok(hip.hipEventSynchronize(e1))
            ms = ctypes.c_float()
            ok(hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1))
            hip.hipEventDestroy(e0)
            hip.hipEventDestroy(e1)
        finally:
            hip.hipFree(b)
    finally:
        hip.hipFree(a)
    t = ms.value / reps
    return {"GBs": 2 * nbytes / (t * 1e-3) / 1e9, "ms": t, "bytes_each_way": nbytes,
            "how": "hipMemcpy device to device of one step's state, average of 10 after 3 warm-up"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_rate(n, steps, threads, runs=5):
    """Median over runs (after a warm-up step) of the oracle's cell-updates/s."""
    import oracle
    from set_amd import mesh
    g = mesh.grid_3d(n, n, n)
    fm = mesh.face_mesh(g)
    W = mesh.heliosphere_state(g)
    s = oracle.Solver(fm, W, oracle.cartesian_config(), nthreads=threads)
    s.dt_next = s.dt_scan()
    s.run(1)  # warm-up
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        st = s.run(steps)
        rates.append(g.n_cell * steps / (time.perf_counter() - t0))
        assert st == 0
    return sorted(rates)[len(rates) // 2]


def cpu_baseline(n, steps, threads):
    v = cpu_rate(n, steps, threads)
    v1 = cpu_rate(max(32, n // 2), max(1, steps // 2), 1)
    return {"value": v, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "single_core_value": v1, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"oracle (C restatement of chlld + face/cell passes) on the same "
                      f"heliosphere set-up at {n}^3, median of 5 runs of {steps} forward-Euler "
                      f"steps on {threads} OpenMP threads; single core: {max(32, n // 2)}^3, "
                      f"{max(1, steps // 2)} steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--forms", default="fp64,mixed,heun,batch64,serving,small,dropin,facelist,ena")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spin-s", type=float, default=0.3,
                    help="untimed device spin-up (seconds of FP64 steps) after the warm-up steps")
    ap.add_argument("--error-steps", type=int, default=200,
                    help="steps of the untimed FP64 and mixed runs behind the precision error table")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-max-steps", type=int, default=5000)
    ap.add_argument("--dropin-mesh", default="2000x1000")
    ap.add_argument("--dropin-steps", type=int, default=20)
    ap.add_argument("--ena-mesh", default="400x200")
    ap.add_argument("--ena-speeds", type=int, default=1000)
    ap.add_argument("--no-copy-ceiling", action="store_true",
                    help="skip the device-copy ceiling measurement")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": a.numel() * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from set_amd import mesh
    n = args.n
    g = mesh.grid_3d(n, n, n)
    W = mesh.heliosphere_state(g)
    forms = {}
    wanted = args.forms.split(",")
    # spin-up: about 0.3 s of device work before the first timed step, the
    # same step count for every form (from the FP64 warm-up's rate)
    spin = 0
    if args.spin_s > 0:
        probe = run_form(g, W, 0, 0, 4, args.warmup)
        spin = min(4000, int(args.spin_s / (probe["ms_per_step"] * 1e-3)))
    fp = run_form(g, W, 0, 0, args.steps, args.warmup, spin)
    forms["fp64_euler"] = fp
    if "mixed" in wanted:
        mx = run_form(g, W, 1, 0, args.steps, args.warmup, spin)
        forms["mixed_euler"] = mx
        mx["speedup_vs_fp64"] = mx["value"] / fp["value"]
        mx["kernel_speedup_vs_fp64"] = fp["kernel_ms_per_launch"] / mx["kernel_ms_per_launch"]
        # the precision form's error against FP64 after the config's 200 steps
        mx["error_vs_fp64"] = precision_error(g, W, args.error_steps)
    if "heun" in wanted:
        forms["fp64_heun"] = run_form(g, W, 0, 1, args.steps, args.warmup, spin)
    for f in forms.values():
        f["hbm_frac"] = f["achieved_GBs"] / PEAK_HBM_GBS
    del W
    for name, fn in (("batch64_fp64", lambda: batch_form(0, args.batch_steps, 20)),
                     ("batch64_mixed", lambda: batch_form(1, args.batch_steps, 20)),
                     ("serving_fp64", lambda: serving_form(args.serve_requests, args.serve_max_steps)),
                     ("small_2d", lambda: small_form()),
                     ("dropin_reference", lambda: dropin_form(*map(int, args.dropin_mesh.split("x")),
                                                              args.dropin_steps)),
                     ("facelist_headline", lambda: facelist_form(n, args.dropin_steps)),
                     ("ena_los", lambda: ena_form(*map(int, args.ena_mesh.split("x")), args.ena_speeds,
                                                  cpu=not args.no_cpu))):
        if name.split("_")[0] in wanted:
            try:
                forms[name] = fn()
            except Exception as e:  # a failing extra form must not lose the headline line
                forms[name] = {"error": repr(e)}
    flf = forms.get("facelist_headline", {})
    if "exact" in flf:
        flf["fused_speedup_vs_exact"] = fp["value"] / flf["exact"]["value"]
        flf["fused_speedup_vs_fast"] = fp["value"] / flf["fast"]["value"]

    traffic = from_profiles(f"fused_fp64_euler_{n}cube", "traffic_bytes_per_launch")
    achieved = fp["achieved_GBs"]
    # achieved: 128 algorithmic bytes per cell-update x cells per launch over
    # the kernel's average launch duration, from HIP events around the timed
    # steps themselves; frac_wall: the same bytes over the timed region's
    # wall clock (launch gaps included).  traffic: HBM bytes per launch from
    # the committed rocprofv3 FETCH_SIZE / WRITE_SIZE profile of this kernel
    # at this size (a profile-derived figure, not measured by this run).
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                "frac_wall": fp["achieved_GBs_wall"] / PEAK_HBM_GBS,
                "kernel": "k_fused<double,16,16,3,Euler> (one launch per step)",
                "traffic_source": f"profiles/fused_fp64_euler_{n}cube.txt" if traffic else None}
    # this box's plain copy rate over the same bytes: the practical ceiling
    if not args.no_copy_ceiling:
        try:
            cc = copy_ceiling(g.n_cell * 64)
            roofline["copy_ceiling"] = {**cc, "frac_of_peak": cc["GBs"] / PEAK_HBM_GBS,
                                        "kernel_frac_of_copy": achieved / cc["GBs"]}
        except Exception as e:  # the headline line must not depend on it
            roofline["copy_ceiling"] = {"error": repr(e)}
    # FP64 flops of the same kernel (VALU counters of the committed profile)
    fl = from_profiles(f"fused_fp64_euler_{n}cube", "fp64_flops_per_cell_update")
    if fl:
        tf = fl * g.n_cell / (fp["kernel_ms_per_launch"] * 1e-3) / 1e12
        fp["compute"] = {"flops_per_cell_update": fl, "achieved": tf, "peak": PEAK_FP64_TFLOPS,
                         "unit": "TFLOP/s (FP64 vector)", "frac": tf / PEAK_FP64_TFLOPS,
                         "flops_source": f"profiles/fused_fp64_euler_{n}cube.txt (VALU counters)"}
        # the VALU-issue ceiling that binds this kernel before HBM does:
        # 1024 SIMDs x 2.4 GHz / 4 cycles per wave instruction over the
        # kernel's wave instructions per cell-update (VALU counter / 64)
        vi = from_profiles(f"fused_fp64_euler_{n}cube", "valu_instructions_per_cell_update")
        if vi:
            ceil = 1024 * 2.4e9 / 4 / (vi / 64.0)
            fp["compute"]["valu_issue_ceiling"] = {
                "valu_instructions_per_cell_update": vi, "cell_updates_per_s": ceil,
                "hbm_frac_at_ceiling": ceil * 128 / (PEAK_HBM_GBS * 1e9),
                "kernel_frac_of_ceiling": g.n_cell / (fp["kernel_ms_per_launch"] * 1e-3) / ceil}
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps, threads)
    line = {
        "metric": METRIC, "value": fp["value"], "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": fp["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded heliosphere initial state)",
        "config": {"workload": f"heliosphere 3-D {n}^3 ideal MHD, first-order HLLD, forward "
                               "Euler, CFL 0.4, box [-1,1]^3", "cells": g.n_cell,
                   "bytes_per_cell_update": 128, "parallelism": "single GPU",
                   "spin_up_steps": spin},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
