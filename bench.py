"""Benchmark: affected-vertex updates/s of the plain one-graph GPU loop,
with the batched and served forms against the plain loop beside it.

Workload (one "step"): the full 10-batch DynLP update stream of a
500,000-vertex, average-degree-7 synthetic graph (generator seed 42, batch
seed 42+12345, tau = mean edge weight, delta 1e-4, maxIterations 100000),
replayed from the pre-batch state on device-resident inputs.  The graph and
batch lists are uploaded before the timed region; the state reset (device
fills) is part of each step.

value = U / step time, U = sum over batches and iterations of the frontier
size (the reference's numAffected), identical for every implementation that
matches the reference's results.

The default line also carries two comparison legs (skip with --no-batched /
--no-serving): "batched" -- 256 graphs (2,000-200,000 vertices, degree
3/5/7, 10 batches each) through the batched runner against the plain loop
over the same graphs, labels compared bit for bit; "serving" -- 64 resident
tenants (10,000-5,000,000 vertices) and 128 update batches all arriving at
t = 0, served with coalesced steps against one batch at a time through the
plain path, every tenant's labels compared bit for bit.

The process runs with GPU_MAX_HW_QUEUES=24 (--hw-queues) so that the
serving legs' hand-back lanes each get a hardware queue of their own.

Prints ONE JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def _hw_queues(argv, default=24):
    """--hw-queues N from the command line (read before argparse runs: the
    HIP runtime takes GPU_MAX_HW_QUEUES when it starts)."""
    for i, arg in enumerate(argv):
        if arg == "--hw-queues" and i + 1 < len(argv):
            return int(argv[i + 1])
        if arg.startswith("--hw-queues="):
            return int(arg.split("=", 1)[1])
    return default


# Serving finishes handed-back graphs on host threads with a stream each (16
# by default) beside the serving loop's stream.  HIP maps a process's
# streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default), and kernels
# of streams that share a queue run one after another, so a lane's long
# LDS-resident launch holds up whatever is queued behind it and at most four
# streams run at once.  24 queues give each of these streams its own.  Set
# here, before anything starts the HIP runtime.
HW_QUEUES = _hw_queues(sys.argv)
if HW_QUEUES > 0:
    os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def propagation_bytes(updates: int, edges: int, changed_edges: int, appended: int) -> int:
    """Algorithmic bytes of propagation iterations (4-byte ids, weights,
    labels): per frontier vertex 20 (queue id 4, row start 4, row length 4,
    F_old[u] 4, F_new[u] 4) + 12 per row entry (column 4, weight 4,
    F_old[v] 4); per changed vertex 4 per stamp update of its row entries and
    itself; 4 per next-frontier append."""
    return 20 * updates + 12 * edges + 4 * changed_edges + 4 * appended


def engine_rooflines(prof: dict, updates: int) -> dict:
    """Per propagation kernel and step: launches, event-timed kernel time,
    mean launch time and algorithmic GB/s (the bytes above over that time).
    Multi-block launches are timed per chunk of back-to-back launches, so
    their mean includes the dispatch gaps inside a chunk."""
    local_u = prof["localUpdates"]
    eng = {
        "k_propagate": (prof["launches"], prof["kernelMs"],
                        propagation_bytes(updates - local_u, prof["edges"],
                                          prof["changedEdges"], prof["appended"])),
        "k_propagate_local": (prof["localLaunches"], prof["localKernelMs"],
                              propagation_bytes(local_u, prof["localEdges"],
                                                prof["localChangedEdges"],
                                                prof["localAppended"])),
    }
    out = {}
    for name, (n, ms, byts) in eng.items():
        if n == 0:
            continue
        per_launch = byts / n
        avg_s = ms / 1000.0 / n
        out[name] = {"launches_per_step": n, "kernel_ms_per_step": ms,
                     "avg_launch_us": avg_s * 1e6,
                     "algorithmic_bytes_per_launch": per_launch,
                     "achieved_gbs": per_launch / avg_s / 1e9 if avg_s > 0 else 0.0}
    return out


def random_access_ceilings(array_ints: int, accesses: int = 1 << 24) -> dict:
    """Random 4-byte accesses per second the GPU sustains over an array of
    the workload's per-vertex size (dlp_random_access_rates)."""
    import ctypes as C
    from labelpropcuda_amd import _lib
    rates = (C.c_double * 3)()
    _lib.check(_lib.load().dlp_random_access_rates(array_ints, accesses, rates))
    return {"array_ints": array_ints, "atomic_max_per_s": rates[0], "load_per_s": rates[1],
            "store_per_s": rates[2]}


def grid_sync_costs(blocks=(64, 512), rounds: int = 2000) -> list:
    """Per-round cost of a device-wide sync between dependent rounds of work
    (dlp_grid_sync_cost): a launch boundary against an in-kernel barrier of
    one counter or of per-work-group flags, at the multi-block kernel's
    smallest grid (512 work-groups) and a small one."""
    import ctypes as C
    from labelpropcuda_amd import _lib
    out = []
    for nb in blocks:
        us, ok = (C.c_double * 3)(), (C.c_int * 3)()
        _lib.check(_lib.load().dlp_grid_sync_cost(nb, rounds, us, ok))
        out.append({"blocks": nb, "rounds": rounds, "launch_us": us[0],
                    "counter_barrier_us": us[1], "flag_barrier_us": us[2],
                    "exchange_ok": bool(ok[0] and ok[1] and ok[2])})
    return out


def random_access_floor_us(prof: dict, updates: int, launches: float, ra: dict) -> float:
    """Time per multi-block launch if its random accesses ran at the measured
    rates back to back: per vertex update the row start, row length and old
    label loads and the new-label store, per row entry the neighbour-label
    load, per claim (a changed vertex's entries and itself) one returning
    atomicMax."""
    u = updates - prof["localUpdates"]
    loads = 3.0 * u + prof["edges"]
    t = loads / ra["load_per_s"] + u / ra["store_per_s"] + prof["changedEdges"] / ra["atomic_max_per_s"]
    return 1e6 * t / max(launches, 1.0)


def add_random_access(engines: dict, prof: dict, updates: int, array_ints: int,
                      out: dict) -> dict:
    """Adds the multi-block kernel's random-access floor (and its ratio to
    the measured launch time) and records the measured rates in out."""
    ra = random_access_ceilings(array_ints)
    out["random_access"] = ra
    k = engines.get("k_propagate")
    if k:
        floor = random_access_floor_us(prof, updates, k["launches_per_step"], ra)
        k["random_access_floor_us"] = floor
        k["random_access_frac"] = floor / k["avg_launch_us"] if k["avg_launch_us"] else None
    return engines


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None when absent)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_c2.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration(mib: int) -> dict:
    """One device-to-device copy of `mib` MiB with the library's own copy
    kernel (k_copy_vec16): reads and writes exactly that many bytes."""
    import ctypes as C
    from labelpropcuda_amd import _lib
    lib = _lib.load()
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    _lib.check(lib.dlp_device_malloc(C.byref(a), n))
    _lib.check(lib.dlp_device_malloc(C.byref(b), n))
    try:
        t0 = time.perf_counter()
        _lib.check(lib.dlp_copy_device(b, a, n))
        dt = time.perf_counter() - t0
    finally:
        lib.dlp_device_free(a)
        lib.dlp_device_free(b)
    return {"bytes": n, "wall_s": dt}


def cpu_baseline(inputs, budget_s: float = 20.0, threads: int = 1):
    """The CPU oracle on the same stream, batch by batch, stopping once the
    time budget is spent; returns updates/s over the batches run.  threads
    > 1 runs its frontier-parallel propagation (same results)."""
    from oracle.oracle import Oracle, set_threads
    set_threads(threads)
    try:
        orc = Oracle(inputs)
        done, upd, t = 0, 0, 0.0
        for b in range(inputs.numBatches):
            t0 = time.perf_counter()
            orc.run_batch(b)
            t += time.perf_counter() - t0
            upd += orc.stats()["updates"]
            done += 1
            if t > budget_s:
                break
    finally:
        set_threads(1)
    return upd / t, done, t


def host_cpu() -> str:
    """CPU model and logical CPU count of this host (the baseline's cores)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


def cpu_baselines(inputs, runs: int = 3, budget_s: float = 10.0) -> dict:
    """The oracle on all host cores (at most 16, the box's share) as
    `cpu_baseline`, and single-threaded beside it: the median of `runs`
    runs, each over the stream's first batches until `budget_s` is spent."""
    cores = max(1, min(16, os.cpu_count() or 1))
    out = {}
    for key, th, n in (("cpu_baseline", cores, runs), ("cpu_baseline_1thread", 1, 1)):
        res = sorted((cpu_baseline(inputs, budget_s, threads=th) for _ in range(n)),
                     key=lambda r: r[0])
        v, done, t = res[len(res) // 2]
        kind = "frontier-parallel, OpenMP" if th > 1 else "single thread"
        out[key] = {"value": v, "unit": "updates/s", "cores": th, "kind": "port",
                    "sample": f"median of {n} runs, each over the first batches of the same "
                              f"stream until {budget_s:.0f} s are spent (median run: {done} of "
                              f"{inputs.numBatches} batches, {t:.1f} s), CPU oracle ({kind}, "
                              "gcc -O3 -ffp-contract=off)",
                    "runs": [r[0] for r in res], "host": host_cpu()}
    return out


def batched_leg(args, progress, autotune: bool) -> dict:
    """The batched workload (256 graphs, vertices log-uniform in [2,000,
    200,000], degree in {3,5,7}, 10 batches each): the plain loop, one graph
    at a time, each graph's full stream (one pass, timed), then the batched
    runner (one propagation launch per batch index, HIP-graph replay) over
    the same graphs; labels compared bit for bit."""
    import labelpropcuda_amd as L
    t0 = time.perf_counter()
    graphs = L.batched_workload(args.graphs, args.batches, 42)
    gen_s = time.perf_counter() - t0
    progress(f"generated {len(graphs)} graphs in {gen_s:.1f} s")
    # plain loop: one graph at a time, each with its own context (the first
    # graph is run once untimed to load the kernels)
    plain_ms, plain_upd, plain_lat, plain_labels = 0.0, 0, [], []
    for gi, inp in enumerate(graphs):
        with L.Context(inp, L.make_config(inp)) as ctx:
            if gi == 0:
                ctx.run_all()
                ctx.synchronize()
            t1 = time.perf_counter()
            ctx.run_all()
            ctx.synchronize()
            plain_ms += 1000.0 * (time.perf_counter() - t1)
            if gi % 32 == 31:
                progress(f"plain loop: {gi + 1} graphs, {plain_ms / 1000.0:.1f} s of updates")
            st = [ctx.stats(b) for b in range(inp.numBatches)]
            plain_upd += sum(x["updates"] for x in st)
            plain_lat.extend(x["timeMs"] for x in st)
            plain_labels.append(ctx.labels())

    m = L.Multi(graphs)
    progress("multi-graph runner created")
    trials = []
    if autotune:
        # work-group shape x hand-back frontier, one step each; a setting
        # whose capturing step is already 1.5x the best so far is dropped
        for shape in (3, 1):
            for heavy in args.heavy_sweep:
                m.set_tuning(localShape=shape, heavyFrontier=heavy, useGraphs=1)
                t1 = time.perf_counter()
                m.run_all()  # captures the batch graphs
                m.synchronize()
                first = time.perf_counter() - t1
                progress(f"autotuner: shape {shape} hand-back {heavy}: capturing step "
                         f"{1000.0 * first:.1f} ms")
                if trials and first > 1.5 * min(trials)[0]:
                    continue
                t1 = time.perf_counter()
                m.run_all()
                m.synchronize()
                trials.append((time.perf_counter() - t1, shape, heavy))
                progress(f"autotuner: shape {shape} hand-back {heavy}: "
                         f"{1000.0 * trials[-1][0]:.1f} ms")
        best = min(trials)
        m.set_tuning(localShape=best[1], heavyFrontier=best[2], useGraphs=1)
    for _ in range(max(args.batched_warmup, 1)):
        m.run_all()  # the first step captures the batch graphs
    m.synchronize()
    progress("batched runner warmed up")
    lat, upd = [], None
    t1 = time.perf_counter()
    for _ in range(args.batched_steps):
        m.run_all()
        m.synchronize()
        u = sum(m.stats(g, b)["updates"] for g, inp in enumerate(graphs)
                for b in range(inp.numBatches))
        assert upd is None or u == upd
        upd = u
        lat.extend(m.batch_time_ms(b) for b in range(args.batches))
    wall = time.perf_counter() - t1
    same = all(np.array_equal(m.labels(g).view(np.int32), plain_labels[g].view(np.int32))
               for g in range(len(graphs)))
    tuning = m.tuning()
    m.close()
    assert upd == plain_upd, "batched and plain update counts differ"
    value = upd / (wall / args.batched_steps)
    plain_value = plain_upd / (plain_ms / 1000.0)
    progress(f"batched {value / 1e6:.1f} M updates/s, plain {plain_value / 1e6:.1f} M, "
             f"labels identical: {same}")
    out = {
        "workload": f"{len(graphs)} graphs, vertices log-uniform in [2000, 200000], "
                    f"degree in {{3,5,7}}, {args.batches} batches (90/1/9) each, graph g seed 42+g",
        "graphs": len(graphs), "vertices_total": int(sum(g.numNodes for g in graphs)),
        "updates_per_step": upd, "generation_s": gen_s,
        "value": value, "unit": "updates/s", "steps": args.batched_steps,
        "ms_per_step": 1000.0 * wall / args.batched_steps,
        "batch_latency_ms": {"p50": float(np.percentile(lat, 50)),
                             "p99": float(np.percentile(lat, 99))},
        "plain_loop": {"value": plain_value, "ms_per_step": plain_ms,
                       "batch_latency_ms": {"p50": float(np.percentile(plain_lat, 50)),
                                            "p99": float(np.percentile(plain_lat, 99))}},
        "speedup_vs_plain": value / plain_value,
        "labels_identical_to_plain": bool(same),
        "tuning": tuning,
    }
    if trials:
        out["autotuner"] = [{"shape": s_, "heavyFrontier": h_, "ms": 1000.0 * t_}
                            for t_, s_, h_ in trials]
    return out


def batched_main(args) -> int:
    """Batched form on its own line (with the in-bench shape sweep)."""
    t0 = time.perf_counter()

    def progress(msg):
        print(f"[bench batched {time.perf_counter() - t0:7.1f} s] {msg}", file=sys.stderr,
              flush=True)

    args.batched_steps = args.steps
    args.batched_warmup = args.warmup
    leg = batched_leg(args, progress, autotune=True)
    out = {
        "metric": f"affected-vertex updates/s (batched, {leg['graphs']} graphs)",
        "value": leg["value"],
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": leg["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (reference generator, graph g seed 42+g)",
        "config": {"workload": leg["workload"], "graphs": leg["graphs"],
                   "vertices_total": leg["vertices_total"],
                   "updates_per_step": leg["updates_per_step"],
                   "generation_s": leg["generation_s"]},
        "batch_latency_ms": leg["batch_latency_ms"],
        "plain_loop": leg["plain_loop"],
        "speedup_vs_plain": leg["speedup_vs_plain"],
        "labels_identical_to_plain": leg["labels_identical_to_plain"],
        "autotuner": leg.get("autotuner"),
        "tuning": leg["tuning"],
    }
    print(json.dumps(out))
    return 0


def serving_setup(args, progress):
    """The serving workload: tenant graphs (vertices log-uniform in [10,000,
    max], degree 5, tenant t seed 1042+t) and the requests (tenant, size
    log-uniform in [100, 100,000]), from numpy.random.default_rng(1042)."""
    import labelpropcuda_amd as L
    rng = np.random.default_rng(1042)
    T, R = args.tenants, args.requests
    sizesN = np.rint(np.exp(rng.uniform(np.log(1e4), np.log(args.max_tenant_nodes), T)))
    reqTenant = rng.integers(0, T, R)
    reqSize = np.rint(np.exp(rng.uniform(np.log(100), np.log(1e5), R))).astype(int)
    tenants = []
    for t in range(T):
        sizes = [0] + [int(k) for k in reqSize[reqTenant == t]]
        tenants.append(L.generate_stream(int(sizesN[t]), 5, sizes, 1042 + t))
        if t % 16 == 15:
            progress(f"generated {t + 1} tenants")
    progress(f"generated {T} tenants, {int(sizesN.sum())} vertices, {R} requests")
    return rng, tenants, reqTenant


def serve_run(m, tenants, reqTenant, arrivals, plain, args, keep_labels=False):
    """One pass of the requests through dlp_multi_serve from the tenants'
    seed state; returns throughput and per-request latency (and the labels
    of every tenant when asked)."""
    T = len(tenants)
    m.reset()
    m.run_next(list(range(T)))  # seed batches
    m.synchronize()
    reqs = list(zip(reqTenant.tolist(), arrivals.tolist()))
    done, steps = m.serve(reqs, maxSlots=args.slots, hbmBudgetBytes=args.hbm_budget,
                          plain=plain, handBackLanes=args.lanes_back,
                          sliceIterations=args.slice_iterations, sliceMs=args.slice_ms,
                          handBackAfter=args.hand_back_after)
    m.synchronize()
    lat = 1000.0 * (done - arrivals)
    assert np.all(lat >= 0.0), "a request completed before it arrived"
    upd = sum(m.stats(t, b)["updates"] for t in range(T)
              for b in range(1, tenants[t].numBatches))
    pool = m.pool_stats()
    out = {"makespan_s": float(done.max()), "steps": steps, "updates": int(upd),
           "pool_peak_bytes": pool["peak"], "pool_deferrals": pool["deferrals"],
           "updates_per_s": upd / float(done.max()),
           "min_ms": float(lat.min()), "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99))}
    if keep_labels:
        out["_labels"] = [m.labels(t) for t in range(T)]
    return out


def median_serve(m, tenants, reqTenant, arrivals, plain, args, progress, label):
    """Several saturated passes (the makespan is about a second): the one of
    median throughput, with every pass's rate, and the labels of the last."""
    runs = []
    for i in range(max(1, args.repeats)):
        runs.append(serve_run(m, tenants, reqTenant, arrivals, plain, args,
                              keep_labels=(i == max(1, args.repeats) - 1)))
        progress(f"{label}, saturated, run {i + 1}: {runs[-1]['makespan_s']:.2f} s")
    assert len({r["updates"] for r in runs}) == 1
    labels = runs[-1].pop("_labels")
    order = sorted(range(len(runs)), key=lambda i: runs[i]["updates_per_s"])
    med = dict(runs[order[len(runs) // 2]])
    med["runs_updates_per_s"] = [r["updates_per_s"] for r in runs]
    return med, labels


def serving_leg(args, progress, sweep_rates) -> dict:
    """Saturated serving (all requests at t = 0) against the plain path
    (one batch at a time through the single-graph loop), labels of every
    tenant compared bit for bit; optionally a sweep of Poisson arrival
    rates."""
    import labelpropcuda_amd as L
    rng, tenants, reqTenant = serving_setup(args, progress)
    T, R = len(tenants), len(reqTenant)
    m = L.Multi(tenants)
    if args.heavy_frontier:
        m.set_tuning(heavyFrontier=args.heavy_frontier)
    progress("tenants resident")
    zero = np.zeros(R)
    sat_plain, lab_plain = median_serve(m, tenants, reqTenant, zero, True, args, progress,
                                        "plain")
    sat_served, lab_served = median_serve(m, tenants, reqTenant, zero, False, args, progress,
                                          "served")
    assert sat_plain["updates"] == sat_served["updates"]
    same = all(np.array_equal(a.view(np.int32), b.view(np.int32))
               for a, b in zip(lab_plain, lab_served))
    progress(f"served {sat_served['updates_per_s'] / 1e6:.1f} M updates/s, plain "
             f"{sat_plain['updates_per_s'] / 1e6:.1f} M, labels identical: {same}")
    service = sat_plain["makespan_s"] / R  # mean plain service time
    sweep = []
    for f in sweep_rates:
        gaps = rng.exponential(service / f, R)
        arrivals = np.cumsum(gaps) - gaps[0]
        row = {"rate_x_plain_capacity": f, "arrivals_per_s": f / service,
               "batched": serve_run(m, tenants, reqTenant, arrivals, False, args),
               "plain": serve_run(m, tenants, reqTenant, arrivals, True, args)}
        sweep.append(row)
        progress(f"rate {f} x capacity: batched p99 {row['batched']['p99_ms']:.1f} ms, "
                 f"plain p99 {row['plain']['p99_ms']:.1f} ms")
    tuning = m.tuning()
    m.close()
    out = {
        "workload": f"{T} tenants, vertices log-uniform in [10000, "
                    f"{int(args.max_tenant_nodes)}], degree 5, tenant t seed 1042+t; {R} "
                    "batches of log-uniform 100-100000 vertices (90/1/9)",
        "slots": args.slots, "hbm_budget_bytes": args.hbm_budget,
        "slice_iterations": args.slice_iterations, "slice_ms": args.slice_ms,
        "hand_back_after": args.hand_back_after,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "updates": sat_served["updates"],
        "saturated": {"served": sat_served, "plain": sat_plain,
                      "speedup_vs_plain": sat_served["updates_per_s"] /
                      sat_plain["updates_per_s"],
                      "labels_identical_to_plain": bool(same)},
        "tuning": tuning,
    }
    if sweep:
        out["rate_sweep"] = sweep
    return out


def serving_main(args) -> int:
    """Serving form on its own line, with the arrival-rate sweep."""
    t0 = time.perf_counter()

    def progress(msg):
        print(f"[bench serving {time.perf_counter() - t0:7.1f} s] {msg}", file=sys.stderr,
              flush=True)

    leg = serving_leg(args, progress, args.rates)
    sat = leg["saturated"]
    out = {
        "metric": f"affected-vertex updates/s (serving, {args.tenants} tenants, saturated)",
        "value": sat["served"]["updates_per_s"],
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": sat["served"]["steps"],
        "warmup": 0,
        "ms_per_step": 1000.0 * sat["served"]["makespan_s"] / max(sat["served"]["steps"], 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (reference graph generator, tenant t seed 1042+t, streamed batches)",
        "config": {"workload": leg["workload"], "slots": leg["slots"],
                   "hbm_budget_bytes": leg["hbm_budget_bytes"], "updates": leg["updates"],
                   "hand_back_after": leg["hand_back_after"], "hw_queues": leg["hw_queues"]},
        "tuning": leg["tuning"],
        "saturated": sat,
        "rate_sweep": leg.get("rate_sweep", []),
    }
    print(json.dumps(out))
    return 0


def stream_main(args) -> int:
    """One pass over a large stream (C3: 50,000,000 vertices, degree 7, 20
    batches), batch by batch with progress lines, for per-batch latency and
    per-engine propagation time; a second pass collects the work counters.
    Per-kernel GB/s come from rocprofv3 runs of this mode."""
    import labelpropcuda_amd as L
    t0 = time.perf_counter()

    def progress(msg):
        print(f"[bench stream {time.perf_counter() - t0:7.1f} s] {msg}", file=sys.stderr,
              flush=True)

    cache = args.inputs_cache
    if cache and os.path.exists(cache):
        inputs = L.Inputs.load(cache)
        progress(f"loaded {cache}")
    else:
        inputs = L.generate(args.nodes, args.degree, args.batches, 42)
        if cache:
            inputs.save(cache)
    N, B = inputs.numNodes, inputs.numBatches
    assert (N, B) == (args.nodes, args.batches), "inputs cache holds another workload"
    progress(f"generated {N} vertices, {inputs.numEdges} entries, {B} batches")
    cfg = L.make_config(inputs)
    ctx = L.Context(inputs, cfg)
    ctx.set_tuning(args.lanes, args.blocks, args.chunk, args.local_threshold,
                   localShape=args.local_shape, localOwnerAfter=args.local_owner_after)
    progress(f"context resident, {ctx.device_bytes() / 2**30:.2f} GiB, tuning {ctx.tuning()}")
    ctx.set_profiling(1)
    ctx.reset()
    ctx.synchronize()
    stats = []
    tw = time.perf_counter()
    for b in range(B):
        ctx.run_batch(b)
        ctx.synchronize()
        st = ctx.stats(b)
        stats.append(st)
        progress(f"batch {b}: {st['iterations']} iterations, {st['updates']} updates, "
                 f"{st['timeMs']:.1f} ms")
    wall = time.perf_counter() - tw
    timed = ctx.prop_profile()
    updates = int(sum(s["updates"] for s in stats))
    prof = None
    if not args.no_work_counters:
        ctx.set_profiling(2)
        ctx.run_all()
        ctx.synchronize()
        prof = ctx.prop_profile()
        for key in ("launches", "kernelMs", "localLaunches", "localKernelMs"):
            prof[key] = timed[key]
        progress("work counters collected")
    ctx.set_profiling(0)
    tuning = ctx.tuning()
    ctx.close()
    lat = [s["timeMs"] for s in stats]
    out = {
        "metric": "affected-vertex updates/s (plain 1-GPU loop)",
        "value": updates / wall,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": 1000.0 * wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (reference generator, seed 42)",
        "config": {"workload": f"DynLP plain loop, {N} vertices, avg degree {args.degree}, "
                               f"{B} batches (90/1/9), tau=mean weight, delta 1e-4",
                   "vertices": N, "entries": inputs.numEdges, "batches": B,
                   "updates_per_step": updates},
        "batch_latency_ms": {"p50": float(np.percentile(lat, 50)),
                             "p99": float(np.percentile(lat, 99))},
        "tuning": tuning,
        "batches": [{k: s[k] for k in ("iterations", "updates", "frontier0", "numSubEdges",
                                       "numComponents", "timeMs", "accuracy")} for s in stats],
        "propagation_time_ms": {"k_propagate": timed["kernelMs"],
                                "k_propagate_local": timed["localKernelMs"],
                                "launches": timed["launches"],
                                "local_launches": timed["localLaunches"]},
    }
    if prof is not None:
        out["propagation_kernels"] = add_random_access(engine_rooflines(prof, updates), prof,
                                                       updates, N, out)
    if args.calibrate_copy_mib > 0:
        out["copy_calibration"] = copy_calibration(args.calibrate_copy_mib)
    print(json.dumps(out))
    return 0


def irlp_main(args) -> int:
    """IrLP, the reference's full-recompute baseline (--validate), on the
    final state of the same 10-batch stream, beside the DynLP stream time it
    is compared with (the reference's "Speedup vs IrLP")."""
    import labelpropcuda_amd as L
    inputs = L.generate(args.nodes, args.degree, args.batches, 42)
    cfg = L.make_config(inputs)
    N, B = inputs.numNodes, inputs.numBatches
    with L.Context(inputs, cfg) as ctx:
        for _ in range(max(args.warmup, 1) + 1):
            ctx.run_all()
            ctx.synchronize()
        dyn_ms = float(sum(ctx.stats(b)["timeMs"] for b in range(B)))
        dyn_acc = float(ctx.stats(B - 1)["accuracy"])

    # launch shape: lanes per vertex with the shortest loop
    sweep = {}
    for g in (1, 4, 8, 16):
        _, r = L.dynlp.irlp_baseline(inputs, cfg, L.dynlp.irlp_tuning(groupLanes=g))
        sweep[g] = r["loopMs"]
    best = min(sweep, key=sweep.get)
    tuning = L.dynlp.irlp_tuning(groupLanes=best)
    for _ in range(args.warmup):
        L.dynlp.irlp_baseline(inputs, cfg, tuning)
    runs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, r = L.dynlp.irlp_baseline(inputs, cfg, tuning)
        runs.append(r)
    wall = time.perf_counter() - t0
    its = runs[0]["iterations"]
    assert all(r["iterations"] == its for r in runs), "IrLP iteration count changed"
    loop_ms = float(np.mean([r["loopMs"] for r in runs]))
    time_ms = float(np.mean([r["timeMs"] for r in runs]))
    E = runs[0]["activeEdges"]
    per_iter_s = loop_ms / 1000.0 / its
    byts = 20 * N + 12 * E  # per sweep: vertex 20 B, stored entry 12 B
    achieved = byts / per_iter_s / 1e9
    value = N * its / (loop_ms / 1000.0)
    out = {
        "metric": "IrLP full-recompute vertex visits/s (validation baseline)",
        "value": value,
        "unit": "vertex visits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (reference generator, seed 42)",
        "config": {"workload": f"IrLP on the final state of the {B}-batch stream of a "
                               f"{N}-vertex, degree-{args.degree} graph, delta 1e-4",
                   "vertices": N, "active_entries": E, "iterations": its,
                   "group_lanes": best, "lanes_sweep_loop_ms": sweep},
        "irlp": {"loop_ms": loop_ms, "validation_call_ms": time_ms,
                 "per_iteration_us": per_iter_s * 1e6, "accuracy": runs[0]["accuracy"]},
        "dynlp": {"total_batch_ms": dyn_ms, "final_batch_accuracy": dyn_acc,
                  "speedup_vs_irlp": time_ms / dyn_ms if dyn_ms > 0 else None,
                  "speedup_vs_irlp_loop_only": loop_ms / dyn_ms if dyn_ms > 0 else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_irlp",
                     "avg_launch_us": per_iter_s * 1e6,
                     "algorithmic_bytes_per_launch": byts},
    }
    if not args.no_cpu:
        from oracle.oracle import irlp as oracle_irlp, irlp_final_state
        rp, ci, vv, st, gt, lab = irlp_final_state(inputs)
        t1 = time.perf_counter()
        oracle_irlp(rp, ci, vv, lab, st, gt, 1e-4, 5)
        five = time.perf_counter() - t1
        n = int(min(its, max(5, 20.0 / max(five / 5, 1e-9))))
        t1 = time.perf_counter()
        oracle_irlp(rp, ci, vv, lab, st, gt, 1e-4, n)
        dt = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": N * n / dt, "unit": "vertex visits/s", "cores": 1,
                               "kind": "port",
                               "sample": f"first {n} of {its} IrLP iterations on the same "
                                         f"final state, {dt:.1f} s, CPU oracle "
                                         "(single thread)"}
    if args.calibrate_copy_mib > 0:
        out["copy_calibration"] = copy_calibration(args.calibrate_copy_mib)
    print(json.dumps(out))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nodes", type=int, default=500000)
    ap.add_argument("--degree", type=int, default=7)
    ap.add_argument("--batches", type=int, default=10)
    ap.add_argument("--lanes", type=int, default=0, help="propagation lanes per vertex")
    ap.add_argument("--blocks", type=int, default=0, help="propagation blocks per launch")
    ap.add_argument("--chunk", type=int, default=0, help="iterations between host checks")
    ap.add_argument("--local-threshold", type=int, default=0,
                    help="frontier size for the LDS-resident kernel (<0 disables)")
    ap.add_argument("--local-shape", type=int, default=0,
                    help="work-group shape of the LDS-resident kernel (1-based)")
    ap.add_argument("--local-owner-after", type=int, default=0,
                    help="iterations of an LDS-resident pass before owner mode (0 = default, "
                         "-1 = never)")
    ap.add_argument("--mask-owner-max", type=int, default=0,
                    help="owner mode runs in the register-mask LDS kernels on frontiers of at "
                         "most this many vertices (0 = default, -1 = never)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--grid-sync", action="store_true",
                    help="c2: add the device-wide sync costs (launch boundary against "
                         "in-kernel barriers; cooperative launches, so not in the default line)")
    ap.add_argument("--workload", choices=("c2", "batched", "serving", "irlp", "stream"),
                    default="c2",
                    help="c2: one 500,000-vertex graph (default); batched: many graphs, "
                         "one propagation launch per batch index; serving: arrival-driven "
                         "continuous batching over resident tenant graphs; irlp: the "
                         "full-recompute baseline on the final state of the c2 stream; "
                         "stream: one pass over --nodes/--degree/--batches batch by batch "
                         "(C3: --nodes 50000000 --batches 20)")
    ap.add_argument("--inputs-cache", default="",
                    help="stream: .npz of the generated inputs, written on the first run "
                         "and loaded by later ones")
    ap.add_argument("--no-work-counters", action="store_true",
                    help="stream: skip the second pass that collects work counters")
    ap.add_argument("--tenants", type=int, default=64, help="tenant graphs (serving)")
    ap.add_argument("--requests", type=int, default=128, help="update batches (serving)")
    ap.add_argument("--max-tenant-nodes", type=float, default=5e6,
                    help="largest tenant graph (serving)")
    ap.add_argument("--slots", type=int, default=64, help="batches per step (serving)")
    ap.add_argument("--hbm-budget", type=int, default=8 << 30,
                    help="serving: bytes of the scratch pool the running batches draw "
                         "their per-batch scratch from (0 = per-tenant scratch)")
    ap.add_argument("--slice-iterations", type=int, default=0,
                    help="serving: iterations a batch advances per step (0 = library "
                         "default, -1 = to the end)")
    ap.add_argument("--slice-ms", type=float, default=0.0,
                    help="serving: ms of HBM iterations per step (0 = library default, "
                         "< 0 = no limit)")
    ap.add_argument("--rates", type=float, nargs="*", default=[0.5, 1.0, 2.0, 4.0],
                    help="arrival rates as multiples of the plain path's capacity (serving)")
    ap.add_argument("--graphs", type=int, default=256, help="graphs of the batched workload")
    ap.add_argument("--heavy-sweep", type=int, nargs="*", default=[-1, 65536, 16384, 4096],
                    help="batched: hand-back frontier sizes the autotuner tries")
    ap.add_argument("--hw-queues", type=int, default=24,
                    help="GPU_MAX_HW_QUEUES for this process (set before the HIP runtime "
                         "starts; 0 = leave the environment's)")
    ap.add_argument("--hand-back-after", type=int, default=0,
                    help="serving: a batch still running at this iteration goes to a "
                         "hand-back lane (0 = default 1024, -1 = never)")
    ap.add_argument("--lanes-back", type=int, default=0,
                    help="serving: hand-back lanes (0 = library default)")
    ap.add_argument("--repeats", type=int, default=3,
                    help="saturated serving runs per form (the median is reported)")
    ap.add_argument("--heavy-frontier", type=int, default=0,
                    help="serving: frontier size at which a graph leaves the shared launch "
                         "(0 = library default, -1 = never)")
    ap.add_argument("--no-batched", action="store_true",
                    help="c2: skip the batched-workload comparison leg")
    ap.add_argument("--no-serving", action="store_true",
                    help="c2: skip the saturated serving comparison leg")
    ap.add_argument("--batched-steps", type=int, default=1,
                    help="c2: timed steps of the batched leg (each is all 256 graphs' streams)")
    ap.add_argument("--batched-warmup", type=int, default=1,
                    help="c2: untimed steps of the batched leg (the first captures)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the run, copy this many MiB device to device with the "
                         "library's copy kernel (reference point for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1"}))
        return 2

    import labelpropcuda_amd as L
    if L.device_count() < 1:
        print(json.dumps({"error": "no GPU visible"}))
        return 1
    if args.workload == "batched":
        return batched_main(args)
    if args.workload == "serving":
        return serving_main(args)
    if args.workload == "irlp":
        return irlp_main(args)
    if args.workload == "stream":
        return stream_main(args)

    t_start = time.perf_counter()
    print(f"[bench] C2 plain loop: {args.nodes} vertices, {args.steps} steps", file=sys.stderr,
          flush=True)
    inputs = L.generate(args.nodes, args.degree, args.batches, 42)
    cfg = L.make_config(inputs)
    ctx = L.Context(inputs, cfg)
    ctx.set_tuning(args.lanes, args.blocks, args.chunk, args.local_threshold,
                   localShape=args.local_shape, localOwnerAfter=args.local_owner_after,
                   maskOwnerMax=args.mask_owner_max)
    tuning = ctx.tuning()

    for _ in range(args.warmup):
        ctx.run_all()
    ctx.synchronize()

    lat, updates = [], None
    ctx.set_profiling(1)  # events per LDS-kernel launch and per multi-block chunk
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.run_all()
        ctx.synchronize()
        st = [ctx.stats(b) for b in range(inputs.numBatches)]
        lat.extend(s["timeMs"] for s in st)
        u = sum(s["updates"] for s in st)
        assert updates is None or u == updates, "update count changed between steps"
        updates = u
    ctx.synchronize()
    wall = time.perf_counter() - t0
    timed = ctx.prop_profile()

    # the work counters (identical every step) from one more step, outside
    # the timed region
    ctx.set_profiling(2)
    ctx.run_all()
    ctx.synchronize()
    work = ctx.prop_profile()
    ctx.set_profiling(0)
    ctx.close()
    prof = dict(work)
    for key in ("launches", "kernelMs", "localLaunches", "localKernelMs"):
        prof[key] = timed[key] / args.steps  # per step, from the timed region

    ms_per_step = 1000.0 * wall / args.steps
    value = updates / (wall / args.steps)
    engines = engine_rooflines(prof, updates)
    ra_out = {}
    engines = add_random_access(engines, prof, updates, args.nodes, ra_out)
    dominant = max(engines, key=lambda k: engines[k]["kernel_ms_per_step"])
    dom = engines[dominant]
    achieved = dom["achieved_gbs"]

    out = {
        "metric": "affected-vertex updates/s (plain 1-GPU loop)",
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (reference generator, seed 42)",
        "config": {"workload": "DynLP plain loop, 500000 vertices, avg degree 7, 10 batches "
                               "(90/1/9), tau=mean weight, delta 1e-4",
                   "vertices": args.nodes, "avg_degree": args.degree, "batches": args.batches,
                   "updates_per_step": updates},
        "batch_latency_ms": {"p50": float(np.percentile(lat, 50)),
                             "p99": float(np.percentile(lat, 99))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dominant) if args.nodes == 500000 else None,
                     "kernel": dominant,
                     "avg_launch_us": dom["avg_launch_us"],
                     "algorithmic_bytes_per_launch": dom["algorithmic_bytes_per_launch"]},
        "propagation_kernels": engines,
        "random_access": ra_out.get("random_access"),
        "tuning": tuning,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
    }
    if args.grid_sync:
        out["grid_sync"] = grid_sync_costs()
    print(f"[bench] C2 done in {time.perf_counter() - t_start:.1f} s: {value / 1e6:.1f} M "
          "updates/s", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out.update(cpu_baselines(inputs))
        print(f"[bench] CPU baselines done at {time.perf_counter() - t_start:.1f} s",
              file=sys.stderr, flush=True)
    t0 = time.perf_counter()

    def progress(msg):
        print(f"[bench {time.perf_counter() - t0:7.1f} s] {msg}", file=sys.stderr, flush=True)

    # the north-star comparisons beside the plain C2 figures: the batched and
    # the served forms against the plain one-graph loop on their workloads
    if not args.no_batched:
        out["batched"] = batched_leg(args, progress, autotune=False)
    if not args.no_serving:
        # saturated, and Poisson arrivals at twice the plain path's capacity
        out["serving"] = serving_leg(args, progress, [2.0])
    if args.calibrate_copy_mib > 0:
        out["copy_calibration"] = copy_calibration(args.calibrate_copy_mib)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
