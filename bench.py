"""Benchmark: exact triangle count of R-MAT scale 26 on one MI355X.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--scale 26] [--no-extras]

One step = one full count of the resident graph: device orientation
(degree order) + intersection + reduction, starting from the symmetric CSR in
HBM and ending with the uint64 count in HBM (read back after the step).
Prints ONE JSON line.  `value` = undirected edges per second.

Extras (same line): the batched 4096-graph workload (persistent pipeline and
hipGraph replay against one launch sequence per graph) and a Poisson serving
run (graphs/s, p50/p99 latency), plus the CPU baseline -- the reference
algorithm (oracle restatement of xxtric.hpp:60-208) on a time-bounded sample
of the same graph, extrapolated.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# HIP runs a process's streams on GPU_MAX_HW_QUEUES hardware queues (4 by
# default).  The serving run has the context's stream and four slot streams
# (and the default stream): on four queues two slots share one, and a small
# batch in one of them waits behind a large request in the other.  Eight give
# every stream a queue of its own; set before anything initialises the GPU
# (raised to 8 when the environment names fewer).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or "4") < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

METRIC = "edges/s + %HBM roofline, R-MAT s26 (1 GPU); batched graphs/s, p99 latency"
HBM_PEAK_GBS = 8000.0
# rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload; its line
# "traffic_bytes_per_launch: N" gives the measured HBM bytes of one launch.
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "intersection_rmat_s26_traffic.md")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def tuning_summary(entry):
    b = entry["best"]
    return {"chosen": {k: v for k, v in b.items() if k != "ms"}, "ms": b["ms"], "default_ms": entry["default_ms"],
            "settings_timed": len(entry["trials"]), "search_s": entry["search_s"]}


def rmat_workload(ctx, scale, steps, warmup, seed=1, tune=True):
    t0 = time.time()
    g = ctx.rmat(scale, 16, seed, True)
    gen_s = time.time() - t0
    m = g.ne // 2
    log("R-MAT s%d generated on device: nv=%d directed=%d (%.1f s)" % (scale, g.nv, g.ne, gen_s))
    tuning = None
    if tune:  # knobs chosen by measurement on this graph, before the timed steps
        from xxtric_amd import autotune
        tuning = tuning_summary(autotune.tune(ctx, "rmat", scale, reps=2, passes=1, graph=g, cache=None,
                                              log=log))
    for _ in range(warmup):
        tri = g.count()
    ev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        tri = g.count()                      # synchronous: ends with the count on the host
        ev.append(ctx.timing())
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / steps
    med = lambda k: float(np.median([e[k] for e in ev]))  # noqa: E731
    st = g.stats()
    out = {"triangles": tri, "nv": g.nv, "m": m, "ms_per_step": ms_step, "gen_s": gen_s,
           "n_heavy": ev[-1]["n_heavy"], "n_tiles": ev[-1]["n_tiles"], "n_wave_tiles": ev[-1]["n_wave_tiles"],
           "tuning": tuning}
    for k in ("orient_ms", "count_ms", "total_ms", "light_ms", "items_ms", "group_ms"):
        out[k] = med(k)
    out.update(st)
    return g, out


def rank0_slice(ctx, g, triangles, parts=8, steps=3):
    """Rank 0 of a `parts`-way 1-D vertex split at its real shape: the owned rows
    [0, nv/parts) plus the ghost rows of every neighbour, counted for the
    triangles whose minimum vertex is owned (and the split by the middle vertex)."""
    lo, hi = 0, g.nv // parts
    t0 = time.perf_counter()
    s = g.slice(lo, hi)
    slice_s = time.perf_counter() - t0
    owned, local, _ = s.count_owned(lo, hi, total=False)  # warm-up
    ev = []
    for _ in range(steps):
        o, l, _ = s.count_owned(lo, hi, total=False)
        assert (o, l) == (owned, local)
        ev.append(ctx.timing())
    whole = g.count_owned(lo, hi, total=False)[:2]
    ms = float(np.median([e["total_ms"] for e in ev]))
    out = {"owned_range": [lo, hi], "parts": parts, "slice_entries": s.ne, "whole_entries": g.ne,
           "slice_build_s": round(slice_s, 3), "owned": owned, "local": local, "remote": owned - local,
           "share_of_total": round(owned / triangles, 6) if triangles else None,
           "ms": round(ms, 3), "edges_per_s": round((s.ne // 2) / (ms * 1e-3), 1),
           "phases_ms": {k: round(float(np.median([e[k + "_ms"] for e in ev])), 3)
                         for k in ("orient", "light", "group")},
           "equal_on_whole_graph": list(whole) == [owned, local]}
    s.close()
    return out


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    return O


def _timed(fn, reps=1):
    best, out = float("inf"), None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        best = min(best, time.perf_counter() - t0)
    return best, out


def small_configs(ctx, steps=5, cpu=True, cpu_seconds=10.0):
    """The reference's CPU-runnable RGG (n = 100,000, LCG seed, reference radius)
    and R-MAT scale 22: device count time (median of `steps`, graph resident),
    and beside each the reference algorithm on the host cores (the oracle's
    restatement of xxtric.hpp:60-208, timed around the local count as
    xxtric.hpp:58,210-224 times it)."""
    import xxtric_amd
    O = _oracle() if cpu else None
    cores = O.nthreads_default() if cpu else 0
    out = {}
    # rgg_1e5: the reference's radius (graph.hpp:789-791) and test (distance <= r);
    # rgg_1e5_connectivity_radius: r = sqrt(2 ln n / (pi n)) and distance < r
    for name, strict, connect in (("rgg_1e5", False, False), ("rgg_1e5_connectivity_radius", True, True)):
        off, tails = xxtric_amd.rgg_generate(100000, strict=strict, connect=connect)
        g = ctx.graph(off, tails)
        tot, pv = g.count(per_vertex=True)
        ms = []
        for _ in range(steps):
            g.count()
            ms.append(ctx.timing()["total_ms"])
        m = len(tails) // 2
        out[name] = {"nv": g.nv, "directed_entries": g.ne, "triangles": tot,
                     "per_vertex_sum_equals_total": int(pv.sum(dtype=np.uint64)) == tot,
                     "ms": round(float(np.median(ms)), 3),
                     "edges_per_s": round(m / (float(np.median(ms)) * 1e-3), 1)}
        if connect:
            out[name]["radius"] = float(xxtric_amd.rgg_radius_connect(100000))
        g.close()
        if cpu:  # the whole count, one thread and all cores, best of 3
            t1, c1 = _timed(lambda: O.count_reference(off, tails, nthreads=1), 3)
            tn, cn = _timed(lambda: O.count_reference(off, tails, nthreads=cores), 3)
            out[name]["cpu_baseline"] = {
                "kind": "port", "unit": "edges/s", "cores": cores, "value": round(m / tn, 1), "s": round(tn, 4),
                "single_thread_value": round(m / t1, 1), "single_thread_s": round(t1, 4),
                "counts_equal": c1 == cn == tot,
                "sample": "whole graph: the reference algorithm (oracle restatement of xxtric.hpp:60-208), "
                          "best of 3 runs"}
    g = ctx.rmat(22, 16, 1, True)
    ms = []
    for _ in range(steps + 1):
        tot = g.count()
        ms.append(ctx.timing()["total_ms"])
    med = float(np.median(ms[1:]))
    out["rmat_s22"] = {"nv": g.nv, "directed_entries": g.ne, "triangles": tot, "ms": round(med, 3),
                       "edges_per_s": round((g.ne // 2) / (med * 1e-3), 1)}
    if cpu:
        o22, t22 = g.download()
        m22 = len(t22) // 2
        t0 = time.perf_counter()
        done, _, edges = O.count_reference_budget(o22, t22, cpu_seconds, cores)
        el = time.perf_counter() - t0
        frac = edges / max(1, len(t22))
        td, cd = _timed(lambda: O.count_degree_ordered(o22, t22, nthreads=cores))
        out["rmat_s22"]["cpu_baseline"] = {
            "kind": "port", "unit": "edges/s", "cores": cores,
            "value": round(m22 / (el / frac), 1) if frac > 0 else 0.0,
            "cap_s": cpu_seconds, "vertices_done": done, "adjacency_fraction_done": round(frac, 6),
            "degree_ordered_exact_s": round(td, 3), "degree_ordered_exact_edges_per_s": round(m22 / td, 1),
            "degree_ordered_count_equal": cd == tot,
            "sample": "the reference algorithm on vertices 0, 1, 2, ... (labels permuted) until the %.0f s cap, "
                      "extrapolated by adjacency fraction; beside it the oracle's exact degree-ordered counter "
                      "over the whole graph" % cpu_seconds}
    g.close()
    # the same tuples without the seeded label permutation: hubs at low ids
    g = ctx.rmat(22, 16, 1, False)
    ms = []
    for _ in range(steps + 1):
        tot = g.count()
        ms.append(ctx.timing()["total_ms"])
    med = float(np.median(ms[1:]))
    out["rmat_s22_unpermuted"] = {"nv": g.nv, "directed_entries": g.ne, "triangles": tot, "ms": round(med, 3),
                                  "edges_per_s": round((g.ne // 2) / (med * 1e-3), 1)}
    if cpu:
        ou, tu = g.download()
        td, cd = _timed(lambda: O.count_degree_ordered(ou, tu, nthreads=cores))
        out["rmat_s22_unpermuted"]["cpu_baseline"] = {
            "kind": "port", "unit": "edges/s", "cores": cores, "value": round((len(tu) // 2) / td, 1),
            "s": round(td, 3), "counts_equal": cd == tot,
            "sample": "the oracle's exact degree-ordered counter over the whole graph"}
    g.close()
    return out


def batch_cpu_baseline(ctx, n_graphs=4096, seconds=10.0):
    """Summed per-graph time of the reference algorithm (one thread per graph)
    over the 4096-graph mix, from the graphs it finishes in `seconds`,
    extrapolated by their share of the mix's entries."""
    import xxtric_amd
    O = _oracle()
    tot_entries = 0
    spent, done, entries = 0.0, 0, 0
    for i in range(n_graphs):
        n, kind, gseed, thr = xxtric_amd.batch_params(2026, i)
        if kind == 1:  # the mix's G(n,p) graphs come from the device generator (as the batch builds them)
            dg = ctx.er(n, gseed + 3, thr)
            o, t = dg.download()
            dg.close()
        else:
            o, t = xxtric_amd.batch_graph_csr(2026, i)
        dt, _ = _timed(lambda: O.count_reference(o, t, nthreads=1))
        spent += dt
        done += 1
        entries += len(t)
        if spent >= seconds:
            break
    # entries of the whole mix without generating it again: the fixture records them
    fx = os.path.join(ROOT, "tests", "golden", "batch_4096.json")
    if n_graphs == 4096 and os.path.exists(fx):
        with open(fx) as f:
            tot_entries = sum(x[2] for x in json.load(f)["graphs"])
    est = spent * tot_entries / entries if tot_entries and entries else float("nan")
    cores = O.nthreads_default()
    return {"kind": "port", "unit": "graphs/s", "cores": 1, "value": round(n_graphs / est, 1),
            "summed_per_graph_s": round(est, 2), "graphs_timed": done, "entries_timed": entries,
            "all_cores_value": round(cores * n_graphs / est, 1), "all_cores": cores,
            "sample": "the reference algorithm (oracle restatement of xxtric.hpp:60-208), one thread per "
                      "graph, on graphs 0..%d of the mix in a %.0f s budget, extrapolated by their share of "
                      "the mix's entries; all_cores_value: %d graphs counted at once (independent graphs)"
                      % (done - 1, seconds, cores)}


def per_vertex_count(ctx, g, triangles, steps=3):
    """The same count with per-vertex output (each triangle credited to its
    smallest vertex, uint64 per vertex in HBM, then copied to the host)."""
    tot, pv = g.count(per_vertex=True)  # warm-up
    ev = []
    for _ in range(steps):
        t, pv = g.count(per_vertex=True)
        assert t == tot
        ev.append(ctx.timing())
    ms = float(np.median([e["total_ms"] for e in ev]))
    return {"ms": round(ms, 3), "edges_per_s": round((g.ne // 2) / (ms * 1e-3), 1),
            "phases_ms": {k: round(float(np.median([e[k + "_ms"] for e in ev])), 3)
                          for k in ("orient", "light", "group")},
            "sum_equals_total": int(pv.sum(dtype=np.uint64)) == tot == triangles}


def roofline(r):
    # Dominant phase: the heavy-edge intersection, k_count_groups (owners with
    # long out-lists) followed by k_count_waves (owners whose out-list fits a
    # wave's table), timed together by the HIP events around them.  Its
    # algorithmic bytes per count: 4 B per id of each heavy edge's shorter
    # out-list plus 4 B per item id, i.e. 4 * (sum over heavy edges of
    # min(d+(u), d+(v)) + heavy edges).
    achieved = r["group_bytes"] / (r["group_ms"] * 1e-3) / 1e9
    whole = r["alg_bytes"] / (r["total_ms"] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", f.read(), re.M)
        traffic = int(m.group(1)) if m else None
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_count_groups + k_count_waves", "kernel_ms": round(r["group_ms"], 3),
            "alg_bytes_per_launch": r["group_bytes"],
            "whole_count": {"alg_bytes": r["alg_bytes"], "ms": round(r["total_ms"], 3),
                            "achieved": round(whole, 1), "frac": round(whole / HBM_PEAK_GBS, 4)}}


def cpu_baseline(g, seconds):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    off, tails = g.download()
    cores = O.nthreads_default()
    t0 = time.time()
    done, tri, edges = O.count_reference_budget(off, tails, seconds, cores)
    el = time.time() - t0
    nv = len(off) - 1
    m = len(tails) // 2
    # extrapolate by the fraction of the adjacency the completed vertices hold
    frac = edges / max(1, len(tails))
    est_s = el / frac if frac > 0 else float("inf")
    # the same on one thread, a shorter budget
    t1 = time.time()
    _, _, e1 = O.count_reference_budget(off, tails, max(1.0, seconds / 4), 1)
    f1 = e1 / max(1, len(tails))
    one = m / ((time.time() - t1) / f1) if f1 > 0 else 0.0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(m / est_s, 1), "unit": "edges/s", "cores": cores, "kind": "port",
            "single_thread_value": round(one, 1), "nproc": os.cpu_count(), "cpu_model": model,
            "sample": "reference algorithm (xxtric.hpp:60-208, oracle restatement, OpenMP) on "
                      "vertices 0..%d of the same graph (random order: labels are permuted) "
                      "in a %.0f s budget: %d of %d vertices holding %.4f%% of the adjacency; "
                      "time extrapolated by adjacency fraction (hub vertices abandoned at the "
                      "deadline, so this overstates the CPU rate); single_thread_value: the same on "
                      "one thread in a quarter of the budget, over its own vertex prefix"
                      % (done, seconds, done, nv, 100 * frac)}


def default_knobs(ctx):
    """The library's own defaults (which depend on the input's size) in place of
    knobs an earlier workload's tuning set explicitly."""
    ctx.tune(small_cut=-1, group_blocks=4, group_round=2, items_per_tile=-1, tile_run=-1, wave_owner=-1,
             group_hash=-1)


def batch_workload(ctx, n_graphs=4096, runs=5, tune=True):
    import xxtric_amd
    from xxtric_amd import XT_BATCH_HIPGRAPH, XT_BATCH_PER_GRAPH, XT_BATCH_PERSISTENT
    b = xxtric_amd.Batch(ctx, 40_000_000, 700_000_000, n_graphs)
    t0 = time.time()
    b.add_generated(2026, 0, n_graphs)
    add_s = time.time() - t0
    out = {"n_graphs": n_graphs, "build_s": round(add_s, 2)}
    # The per-graph baseline counts each graph as a single graph, with the
    # library's defaults for graphs of its size; the batched modes then run
    # with the knobs tuned on the batch.
    default_knobs(ctx)
    res = {}
    modes = [("per_graph", XT_BATCH_PER_GRAPH, 2)]
    if tune:
        modes.append(("tune", None, 0))
    modes += [("persistent", XT_BATCH_PERSISTENT, runs), ("hipgraph", XT_BATCH_HIPGRAPH, runs)]
    for name, mode, k in modes:
        if name == "tune":  # the same knobs for both batched modes
            from xxtric_amd import autotune
            out["tuning"] = tuning_summary(autotune.tune(ctx, "batch", graphs=n_graphs, reps=3, passes=1, batch=b,
                                                         cache=None, log=log))
            continue
        b.run(mode)  # warm-up (captures the graph once for hipgraph)
        ms = []
        for _ in range(k):
            b.run(mode)
            ms.append(b.timing()["run_ms"])
        res[name] = b.results().copy()
        out[name + "_ms"] = round(float(np.median(ms)), 3)
        out[name + "_graphs_per_s"] = round(n_graphs / (np.median(ms) * 1e-3), 1)
    out["speedup_vs_per_graph"] = round(out["per_graph_ms"] / out["hipgraph_ms"], 1)
    out["results_agree"] = bool((res["per_graph"] == res["persistent"]).all() and
                                (res["persistent"] == res["hipgraph"]).all())
    fx = os.path.join(ROOT, "tests", "golden", "batch_4096.json")
    if n_graphs == 4096 and os.path.exists(fx):  # per-graph totals of the oracle's degree-ordered counter
        with open(fx) as f:
            exp = json.load(f)
        tot = np.array([x[3] for x in exp["graphs"]], np.uint64)  # [n, kind, entries, triangles]
        out["results_equal_oracle_fixture"] = bool(np.array_equal(res["hipgraph"].astype(np.uint64), tot))
    t = b.timing()
    out["vertices"] = t["n_vertices"]
    out["directed_entries"] = t["n_edges"]
    out["triangles_total"] = int(res["hipgraph"].sum())
    # HBM roofline of the batched count: the per-graph algorithmic bytes
    # 8(n+1) + 4|E+| + 4 sum min(d+(u), d+(v)), summed over the graphs, over the
    # hipGraph replay's device time
    st = b.stats()
    ach = st["alg_bytes"] / (out["hipgraph_ms"] * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "alg_bytes": st["alg_bytes"], "achieved": round(ach, 1),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                       "sum_min": st["sum_min"], "ne_plus": st["ne_plus"], "heavy_edges": st["heavy_edges"],
                       "group_bytes": st["group_bytes"]}
    b.close()
    return out


_SERVE_POOL = {}


def _serve_pool(ctx, pool):
    """The request pool: `pool` graphs of the batch mix plus R-MAT s18, s19 and
    s20 (the 1 % large requests), each with its solo count on this GPU."""
    import xxtric_amd
    if pool not in _SERVE_POOL:
        graphs = [xxtric_amd.batch_graph_csr(2026, g) for g in range(pool)]
        for s in (18, 19, 20):
            dg = ctx.rmat(s, 16, 100 + s, True)
            graphs.append(dg.download())
            dg.close()
        graphs = [(np.ascontiguousarray(o, np.int64), np.ascontiguousarray(t, np.int32)) for o, t in graphs]
        solo = np.array([ctx.count_csr(o, t) for o, t in graphs], np.uint64)
        _SERVE_POOL[pool] = (graphs, solo)
    return _SERVE_POOL[pool]


def serve_workload(ctx, rate, n_req=4000, pool=400, seed=7, inflight=4):
    """Poisson arrivals at `rate` graphs/s of the batch mix plus 1 % R-MAT
    s18-s20 requests, submitted by the library's native request generator
    (xt_serve_replay: borrowed submissions at their scheduled times, every
    result checked against the graph's solo count)."""
    import xxtric_amd
    rng = np.random.default_rng(seed)
    graphs, solo = _serve_pool(ctx, pool)
    kinds = rng.integers(0, pool, n_req)
    big = rng.random(n_req) < 0.01
    kinds[big] = pool + rng.integers(0, 3, big.sum())
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req)) * 1e6  # us
    srv = xxtric_amd.Server(ctx, hbm_budget_bytes=16 << 30, max_batch_graphs=4096, max_inflight=inflight)
    base = srv.now_us() + 1000.0
    bad, lat = srv.replay(graphs, solo, kinds, base + arrivals)
    st = srv.stats()
    srv.close()
    ok = lat >= 0
    small, large = lat[ok & (kinds < pool)], lat[ok & (kinds >= pool)]
    pct = lambda v, q: round(float(np.percentile(v, q)), 1) if len(v) else None  # noqa: E731
    return {"rate_graphs_per_s": rate, "requests": n_req, "slots": inflight, "graphs_per_s": round(st["graphs_per_s"], 1),
            "p50_us": round(st["p50_us"], 1), "p99_us": round(st["p99_us"], 1),
            "batches": st["batches"], "isolated_batches": st["isolated_batches"],
            "rejected": st["rejected"], "counts_equal_solo": bad == 0 and st["rejected"] == 0,
            "submitter": "native (xt_serve_replay)",
            "hip_hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
            "worker_s": {k: round(st[k], 4) for k in ("admit_s", "stage_s", "launch_s", "complete_s", "device_s")},
            # the same latencies split by request class: the pool's small graphs and the
            # 1 % R-MAT s18-s20 requests
            "small": {"requests": int(len(small)), "p50_us": pct(small, 50), "p99_us": pct(small, 99)},
            "rmat_s18_s20": {"requests": int(len(large)), "p50_us": pct(large, 50), "p99_us": pct(large, 99)}}


def copy_calibration(ctx, nbytes=1 << 30, reps=3):
    """Plain copy kernel of known size: the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run to HBM bytes."""
    import xxtric_amd
    from xxtric_amd import capi
    src = xxtric_amd.DeviceBuffer(ctx, nbytes)
    dst = xxtric_amd.DeviceBuffer(ctx, nbytes)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx._check(ctx._L.xt_memcpy(ctx.h, dst.ptr, src.ptr, nbytes, capi.XT_D2D_KERNEL))
        ts.append(time.perf_counter() - t0)
    src.close()
    dst.close()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps, "kernel": "k_copy16",
                      "read_KiB_expected": nbytes // 1024, "write_KiB_expected": nbytes // 1024,
                      "best_s": min(ts), "GBps_read_plus_write": 2 * nbytes / min(ts) / 1e9}))


def memory_ceilings(ctx, reps=3):
    """The read paths the counting kernels are bound by, on this GPU: a streaming
    read (the group kernel's lists) and random gathers of 2-, 4- and 8-byte
    elements (the orientation and light passes' neighbour keys and records),
    from arrays the size of R-MAT s26's, and one that fits the Infinity Cache."""
    import ctypes
    out = {"streaming_read": [], "random_gather": []}
    for nbytes in (128 << 20, 8 << 30):
        ms = ctypes.c_double()
        ctx._check(ctx._L.xt_mem_ceiling(ctx.h, 0, nbytes // 16, 16, 0, reps, ctypes.byref(ms)))
        out["streaming_read"].append({"bytes": nbytes, "ms": round(ms.value, 3),
                                      "GBps": round(nbytes / (ms.value * 1e-3) / 1e9, 1)})
    ng = 1 << 31
    for eb, n in ((2, 1 << 26), (4, 1 << 26), (8, 1 << 26), (2, 1 << 22)):
        ms = ctypes.c_double()
        ctx._check(ctx._L.xt_mem_ceiling(ctx.h, 1, n, eb, ng, reps, ctypes.byref(ms)))
        out["random_gather"].append({"elem_bytes": eb, "array_bytes": n * eb, "gathers": ng,
                                     "ms": round(ms.value, 3), "Ggathers_per_s": round(ng / (ms.value * 1e-3) / 1e9, 2)})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=5000.0,
                    help="offered Poisson arrival rate (graphs/s) of the serving line; "
                         "--serve-sweep measures capacity")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB copy kernel used to scale PMC byte counters")
    ap.add_argument("--serve-sweep", action="store_true",
                    help="only run the serving workload at 1k..100k graphs/s offered, 60 s or 100k requests "
                         "per rate, and print one line per rate")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the batched 4096-graph workload")
    ap.add_argument("--ceilings", action="store_true",
                    help="only measure the streaming-read and random-gather ceilings of this GPU")
    ap.add_argument("--per-vertex-only", action="store_true",
                    help="only run the per-vertex count of the R-MAT graph (--scale), library defaults")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default knobs instead of tuning them on each workload first")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the single-device path only "
              "(multi-GPU sharding is a design, not implemented)", file=sys.stderr)
        sys.exit(2)
    import xxtric_amd
    ctx = xxtric_amd.Context(0)
    if args.copy_calibration:
        copy_calibration(ctx)
        return
    if args.ceilings:
        print(json.dumps(memory_ceilings(ctx)), flush=True)
        return
    if args.serve_sweep:
        for rate in (1000, 2000, 5000, 10000, 20000, 50000, 100000):
            print(json.dumps(serve_workload(ctx, float(rate), n_req=min(100000, 60 * rate))), flush=True)
        return
    tune = not args.no_autotune
    if args.per_vertex_only:
        g = ctx.rmat(args.scale, 16, 1, True)
        tri = g.count()
        out = per_vertex_count(ctx, g, tri)
        out.update({"workload": "rmat_s%d_ef16_whole_graph" % args.scale, "triangles": tri})
        print(json.dumps(out), flush=True)
        return
    if args.batch_only:
        print(json.dumps(batch_workload(ctx, tune=tune)), flush=True)
        return
    g, r = rmat_workload(ctx, args.scale, args.steps, args.warmup, tune=tune)
    value = r["m"] / (r["ms_per_step"] * 1e-3)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "edges/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic: seeded R-MAT generated on the device (a=.57 b=.19 c=.19, labels permuted)",
        "config": {"workload": "rmat_s%d_ef16_whole_graph" % args.scale, "scale": args.scale,
                   "edgefactor": 16, "seed": 1, "nv": r["nv"], "m_undirected": r["m"],
                   "orientation": "degree", "triangles": r["triangles"]},
        "roofline": roofline(r),
        "phases_ms": {k: round(r[k + "_ms"], 3) for k in ("orient", "light", "items", "group", "count", "total")},
        "teps_reference_style": round(2 * r["m"] / (r["ms_per_step"] * 1e-3), 1),
        "edges_per_s_device": round(r["m"] / (r["total_ms"] * 1e-3), 1),
        "ne_plus": r["ne_plus"], "sum_min": r["sum_min"], "heavy_edges": r["heavy_edges"],
        "sum_min_heavy": r["sum_min_heavy"], "group_tiles": r["n_tiles"], "wave_tiles": r["n_wave_tiles"],
        "tuning": r["tuning"],
    }
    line["cpu_baseline"] = cpu_baseline(g, args.cpu_seconds)
    if not args.no_extras:
        try:
            line["per_vertex"] = per_vertex_count(ctx, g, r["triangles"])
        except Exception as e:
            line["per_vertex"] = {"error": str(e)}
        try:
            line["rank0_of_8"] = rank0_slice(ctx, g, r["triangles"])
        except Exception as e:
            line["rank0_of_8"] = {"error": str(e)}
    g.close()
    if not args.no_extras:
        default_knobs(ctx)  # the other workloads start from the library defaults
        try:
            line["small_configs"] = small_configs(ctx)
        except Exception as e:
            line["small_configs"] = {"error": str(e)}
        try:
            line["batch_4096"] = batch_workload(ctx, tune=tune)
            line["batch_4096"]["cpu_baseline"] = batch_cpu_baseline(ctx)
        except Exception as e:  # reported, not hidden
            line["batch_4096"] = dict(line.get("batch_4096") or {}, error=str(e))
        try:
            default_knobs(ctx)
            line["serve"] = serve_workload(ctx, args.serve_rate)
        except Exception as e:
            line["serve"] = {"error": str(e)}
        try:  # the same at other offered rates: latency below saturation, capacity above it
            sweep = []
            for rate in (1000.0, 10000.0, 20000.0, 100000.0):
                r = serve_workload(ctx, rate, n_req=max(2000, min(40000, int(rate))))
                sweep.append({k: r[k] for k in ("rate_graphs_per_s", "graphs_per_s", "p50_us", "p99_us",
                                                "counts_equal_solo")})
                sweep[-1]["small_p99_us"] = r["small"]["p99_us"]
            line["serve_sweep"] = sweep
        except Exception as e:
            line["serve_sweep"] = {"error": str(e)}
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
