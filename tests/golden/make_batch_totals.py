"""Writes tests/golden/batch_4096.json: the expected triangle count of every
graph of the 4096-graph batch mix (base seed 2026, vertex counts log-uniform
in [200, 20000], random geometric graphs at even indices, Erdos-Renyi at odd
ones), computed entirely by the oracle:

  * each graph is rebuilt by the oracle's own generators (or_rgg_lcg, the
    reference's all-pairs RGG with the LCG stream of graph.hpp:745-1374, and
    or_er_csr), from or_batch_params -- not by the library;
  * its count is or_count_degree_ordered, cross-checked on every graph
    against the restatement of the reference's CPU count
    (or_count_reference, xxtric.hpp:60-208) where that finishes quickly.

Run in the build container after `make -C oracle all`:

    python tests/golden/make_batch_totals.py

The fixture holds, per graph: n, kind, directed entries and the count.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_py as O  # noqa: E402

BASE_SEED = 2026
COUNT = 4096
NMIN, NMAX = 200, 20000


def graph(g):
    n, kind, gseed, thr = O.batch_params(BASE_SEED, g, NMIN, NMAX)
    if kind == 0:
        off, tails = O.rgg_lcg(n, seed=(gseed ^ (gseed >> 32)) & 0xFFFFFFFF)
    else:
        off, tails = O.er_csr(n, gseed + 3, thr)
    return n, kind, off, tails


def main():
    out = {"base_seed": BASE_SEED, "nmin": NMIN, "nmax": NMAX, "graphs": []}
    t0 = time.time()
    cross = 0
    for g in range(COUNT):
        n, kind, off, tails = graph(g)
        tri = O.count_degree_ordered(off, tails)
        if g % 16 == 0:  # the reference-shaped count on a spread of graphs
            assert O.count_reference(off, tails) == tri, g
            cross += 1
        out["graphs"].append([n, kind, int(off[-1]), tri])
        if g % 256 == 0:
            print(g, n, kind, int(off[-1]), tri, "%.0fs" % (time.time() - t0), flush=True)
    out["cross_checked_with_reference_count"] = cross
    with open(os.path.join(HERE, "batch_4096.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("entries", sum(x[2] for x in out["graphs"]), "triangles", sum(x[3] for x in out["graphs"]))


if __name__ == "__main__":
    main()
