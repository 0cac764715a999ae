"""Autotuner for the counting pipeline's launch knobs.

Seven knobs decide how the work is split; none changes a result:
  * light_cut      -- edges whose out-lists are both at most this long are
                      intersected in the light pass; longer ones go to the
                      group kernel (LDS-staged list, streamed partner lists)
                      or to wave tiles;
  * wave_owner     -- owners whose out-list is at most this long are counted
                      one wave per tile with a small per-wave table (-2: none);
  * group_round    -- 16-byte list chunks each group-kernel thread keeps in
                      flight per round (1, 2 or 4: more loads in flight against
                      more registers, hence fewer resident workgroups);
  * group_blocks   -- persistent group-kernel workgroups requested per CU;
  * items_per_tile -- heavy edges of one owner per group tile (64, 128, 256);
  * tile_run       -- consecutive tiles a workgroup draws from the work queue
                      at once (an owner's tiles are consecutive, so longer runs
                      keep its staged table more often);
  * group_hash     -- the group kernel's table hash: 1 multiplicative (spreads
                      any labels), 2 the id's low bits (one multiply fewer per
                      streamed id, enough for randomly labelled graphs).
The reference fixes its launch shapes with compile-time macros
(counting.cuh:150-176); here they are context knobs (xt_ctx_tune,
xt_ctx_tune_tiles, xt_ctx_tune_waves, xt_ctx_tune_hash) chosen by measurement.

The search is coordinate descent from the library defaults: each knob in turn
is swept with the others held at the best setting so far, for at most
`passes` rounds or until a round changes nothing.  Every setting is timed
with the library's HIP-event timers (median of `reps` runs) and its results
are checked to be identical to the first setting's.  The trials and the
choice are cached per workload in a JSON file, so later contexts apply the
choice without measuring again.

    python -m xxtric_amd.autotune --workload rmat --scale 26
    python -m xxtric_amd.autotune --workload batch --graphs 4096
"""
import argparse
import json
import os
import time

import numpy as np

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
SPACE = {
    "light_cut": (4, 8, 16, 24, 32, 48, 64, 96),
    "wave_owner": (-2, 64, 128, 256, 512),
    "group_round": (1, 2, 4),
    "group_blocks": (2, 3, 4),
    "items_per_tile": (64, 128, 256),
    "tile_run": (4, 8, 16, 32, 64),
    "group_hash": (1, 2),
}
DEFAULTS = {"light_cut": 32, "wave_owner": 256, "group_round": 2, "group_blocks": 4, "items_per_tile": 256,
            "tile_run": 16, "group_hash": 1}
BATCH_DEFAULTS = dict(DEFAULTS, light_cut=8, wave_owner=64)


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
    b.run(XT_BATCH_HIPGRAPH)  # capture (again when a knob changed)
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


def apply(ctx, knobs):
    ctx.tune(small_cut=knobs["light_cut"], group_blocks=knobs["group_blocks"],
             group_round=knobs["group_round"], items_per_tile=knobs["items_per_tile"],
             tile_run=knobs["tile_run"], wave_owner=knobs.get("wave_owner", -1),
             group_hash=knobs.get("group_hash", -1))


def search(ctx, runner, start, space=SPACE, passes=2, log=print, label=""):
    """Coordinate descent over `space` from `start`; runner() -> (ms, results)."""
    seen = {}
    ref = [None]

    def evaluate(knobs):
        key = tuple(knobs[k] for k in space)
        if key not in seen:
            apply(ctx, knobs)
            ms, res = runner()
            if ref[0] is None:
                ref[0] = res
            elif not np.all(np.asarray(res) == np.asarray(ref[0])):
                raise AssertionError("results differ at %s" % knobs)
            seen[key] = dict(knobs, ms=round(ms, 3))
            log("autotune %s %s: %.3f ms" % (label, " ".join("%s=%d" % (k, knobs[k]) for k in space), ms))
        return seen[key]["ms"]

    best = dict(start)
    best_ms = evaluate(best)
    for _ in range(passes):
        changed = False
        for name, values in space.items():
            for v in values:
                if v == best[name]:
                    continue
                cand = dict(best, **{name: v})
                ms = evaluate(cand)
                if ms < best_ms:
                    best, best_ms, changed = cand, ms, True
        if not changed:
            break
    apply(ctx, best)
    return dict(best, ms=best_ms), list(seen.values())


def tune(ctx, workload="rmat", scale=22, graphs=1024, reps=3, passes=2, cache=DEFAULT_CACHE, log=print,
         graph=None, batch=None):
    """Tune on a sample workload (or the given resident graph / batch), apply
    and cache the fastest setting.  Returns {"best", "trials", ...}."""
    import xxtric_amd
    if workload == "rmat":
        g = graph if graph is not None else ctx.rmat(scale, 16, 1, True)
        runner = lambda: _time_graph(ctx, g, reps)  # noqa: E731
        key = "rmat_s%d" % scale
        start = DEFAULTS
    elif workload == "batch":
        b = batch
        if b is None:
            b = xxtric_amd.Batch(ctx, 40_000_000, 400_000_000, graphs)
            b.add_generated(2026, 0, graphs)
        runner = lambda: _time_batch(b, reps)  # noqa: E731
        key = "batch_%d" % graphs
        start = BATCH_DEFAULTS
    else:
        raise ValueError(workload)
    t0 = time.perf_counter()
    best, trials = search(ctx, runner, start, passes=passes, log=log, label=key)
    default_ms = next(t["ms"] for t in trials if all(t[k] == start[k] for k in SPACE))
    entry = {"best": best, "default_ms": default_ms, "trials": trials,
             "search_s": round(time.perf_counter() - t0, 2), "time": time.strftime("%Y-%m-%d %H:%M:%S")}
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
    b = dict(DEFAULTS, **{k: v for k, v in entry["best"].items() if k in SPACE})
    apply(ctx, b)
    return b


def main():
    import xxtric_amd
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=("rmat", "batch"), default="rmat")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--graphs", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    a = ap.parse_args()
    with xxtric_amd.Context(0) as ctx:
        e = tune(ctx, a.workload, a.scale, a.graphs, a.reps, a.passes, cache=a.cache)
        print(json.dumps(e["best"]))


if __name__ == "__main__":
    main()
