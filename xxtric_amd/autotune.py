"""Autotuner for the counting pipeline's launch knobs.

Three knobs decide how work is split (xt_ctx_tune):
  * light_cut    -- edges whose out-lists are both at most this long are
                    intersected one per thread; longer ones go to the group
                    kernel (LDS-staged list, streamed partner lists);
  * group_blocks -- persistent group-kernel workgroups per CU (each holds a
                    32 KB LDS hash table, so at most four fit a CU);
  * group_round  -- 16-byte list chunks each group-kernel thread keeps in
                    flight per round (1, 2 or 4: more loads in flight against
                    more registers, hence fewer resident workgroups).
The tuner times every combination on a sample workload with the library's own
HIP-event timers (median of `reps` counts), keeps the fastest, and caches the
choice per device and workload in a JSON file so later contexts apply it
without re-measuring.  Counts are checked to be identical across all settings:
a knob may change speed, never the result.

    python -m xxtric_amd.autotune --workload rmat --scale 22
    python -m xxtric_amd.autotune --workload batch --graphs 1024
"""
import argparse
import json
import os
import time

import numpy as np

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
LIGHT_CUTS = (8, 16, 32, 64, 128)
GROUP_BLOCKS = (2, 3, 4)
GROUP_ROUNDS = (1, 2, 4)


def _time_graph(ctx, g, reps):
    out = []
    tri = None
    for _ in range(reps):
        t = g.count()
        if tri is None:
            tri = t
        elif t != tri:
            raise AssertionError("count changed between runs: %d vs %d" % (tri, t))
        out.append(ctx.timing()["total_ms"])
    return float(np.median(out)), tri


def _time_batch(b, reps):
    from . import XT_BATCH_HIPGRAPH
    b.run(XT_BATCH_HIPGRAPH)  # capture
    ms = []
    res = None
    for _ in range(reps):
        b.run(XT_BATCH_HIPGRAPH)
        ms.append(b.timing()["run_ms"])
        r = b.results().copy()
        if res is None:
            res = r
        elif not (r == res).all():
            raise AssertionError("batch results changed between runs")
    return float(np.median(ms)), res


def tune(ctx, workload="rmat", scale=22, graphs=1024, reps=3, cuts=LIGHT_CUTS,
         blocks=GROUP_BLOCKS, rounds=GROUP_ROUNDS, cache=DEFAULT_CACHE, log=print):
    """Sweep the knobs on a sample workload; apply and cache the fastest."""
    import xxtric_amd
    trials = []
    if workload == "rmat":
        g = ctx.rmat(scale, 16, 1, True)
        runner = lambda: _time_graph(ctx, g, reps)  # noqa: E731
        key = "rmat_s%d" % scale
    elif workload == "batch":
        b = xxtric_amd.Batch(ctx, 40_000_000, 400_000_000, graphs)
        b.add_generated(2026, 0, graphs)
        runner = lambda: _time_batch(b, reps)  # noqa: E731
        key = "batch_%d" % graphs
    else:
        raise ValueError(workload)
    ref = None
    for c in cuts:
        for gb in blocks:
            for gr in rounds:
                ctx.tune(small_cut=c, group_blocks=gb, group_round=gr)
                ms, res = runner()
                if ref is None:
                    ref = res
                same = bool(np.all(np.asarray(res) == np.asarray(ref)))
                if not same:
                    raise AssertionError("results differ at light_cut=%d group_blocks=%d group_round=%d"
                                         % (c, gb, gr))
                trials.append({"light_cut": c, "group_blocks": gb, "group_round": gr,
                               "ms": round(ms, 3)})
                log("autotune %s light_cut=%d group_blocks=%d group_round=%d: %.3f ms"
                    % (key, c, gb, gr, ms))
    best = min(trials, key=lambda t: t["ms"])
    _apply(ctx, best)
    entry = {"best": best, "trials": trials, "time": time.strftime("%Y-%m-%d %H:%M:%S")}
    if cache:
        data = {}
        if os.path.exists(cache):
            with open(cache) as f:
                data = json.load(f)
        data[key] = entry
        with open(cache, "w") as f:
            json.dump(data, f, indent=1)
    return entry


def apply_cached(ctx, key, cache=DEFAULT_CACHE):
    """Apply a cached choice; returns it, or None when the cache has no entry."""
    if not os.path.exists(cache):
        return None
    with open(cache) as f:
        data = json.load(f)
    entry = data.get(key)
    if not entry:
        return None
    b = entry["best"]
    _apply(ctx, b)
    return b


def _apply(ctx, b):
    ctx.tune(small_cut=b["light_cut"], group_blocks=b["group_blocks"],
             group_round=b.get("group_round", 0))


def main():
    import xxtric_amd
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("rmat", "batch"), default="rmat")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--graphs", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    a = ap.parse_args()
    with xxtric_amd.Context(0) as ctx:
        e = tune(ctx, a.workload, a.scale, a.graphs, a.reps, cache=a.cache)
        print(json.dumps(e["best"]))


if __name__ == "__main__":
    main()
