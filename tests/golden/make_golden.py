"""Writes tests/golden/golden.json from runs of the reference's own CPU path.

Run in the build container (where /root/reference exists) after
`make -C oracle all ref`:

    python tests/golden/make_golden.py

What it records (every count comes from the reference binary oracle/_ref/tric_ref,
built from main.cpp/xxtric.hpp/graph.hpp where they lie):
  * rgg_lcg: `tric_ref -n N -l` -> "Number of edges" and "Number of triangles"
    (the reference's seeded LCG random geometric graph, graph.hpp:745-1374).
  * partitions: the vertex bins of the reference's balanced reads (-b, -d),
    from oracle/_ref/ref_partition (our driver over the reference's
    BinaryEdgeList routines, graph.hpp:397-602) on fixture offsets.
    `python tests/golden/make_golden.py --partitions` refreshes only these.
  * rgg_ranks: the RGG the reference's generator builds on P MPI ranks, from
    `mpiexec -n P oracle/_ref/ref_rgg_ranks N FILE` (our driver over
    GenerateRGG, graph.hpp:745-1097): its CSR sha256 and entry count, and the
    triangles `tric_ref -f` counts on it (the reference's own distributed count
    segfaults at P > 1 in this container).  `--ranks` refreshes only these.
  * vite inputs read with `tric_ref -f FILE`: R-MAT and Erdos-Renyi graphs of
    this build's own generators and amazon0302 (cugraph/example-inputs), each
    with a sha256 of its CSR so a test can tell that the same graph is rebuilt.
The fixture holds only inputs' parameters, hashes and the expected outputs.
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_py as O  # noqa: E402

REF = O.REF_BIN
AMAZON_TSV = "/root/reference/cugraph/example-inputs/amazon0302_adj.tsv"


def csr_sha256(off, tails):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(off, np.int64).tobytes())
    h.update(np.ascontiguousarray(tails, np.int64).tobytes())
    return h.hexdigest()


def run_ref(args):
    out = subprocess.run([REF] + args, capture_output=True, text=True, check=True, timeout=3600).stdout
    ne = int(re.search(r"Number of edges: (\d+)", out).group(1))
    nv = int(re.search(r"Number of vertices: (\d+)", out).group(1))
    tri = int(re.search(r"Number of triangles: (-?\d+)", out).group(1))
    return nv, ne, tri


def ref_on_csr(off, tails, tmp, name):
    path = os.path.join(tmp, name + ".bin")
    O.vite_write(path, off, tails)
    nv, ne, tri = run_ref(["-f", path])
    os.unlink(path)
    assert nv == len(off) - 1 and ne == len(tails), (name, nv, ne)
    return tri


def amazon_csr():
    import pandas as pd
    df = pd.read_csv(AMAZON_TSV, sep="\t", header=None, names=["s", "d", "w"], dtype=np.int64)
    s = df["s"].to_numpy()
    d = df["d"].to_numpy()
    nv = int(max(s.max(), d.max())) + 1  # ids are 1..262111; vertex 0 stays isolated
    key = np.unique(s * nv + d)
    src = key // nv
    dst = key % nv
    off = np.zeros(nv + 1, np.int64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off)
    return off, dst.astype(np.int64)


def ref_partition(off, nparts, method, tmp):
    """Bins of the reference's own routines (oracle/_ref/ref_partition) for
    the offsets `off` written as a vite file (the tails do not matter)."""
    path = os.path.join(tmp, "part.bin")
    O.vite_write(path, off, np.zeros(int(off[-1]), np.int64))
    out = subprocess.run([O.REF_PARTITION_BIN, path, str(nparts), str(method)], capture_output=True,
                         text=True, check=True, timeout=3600).stdout
    os.unlink(path)
    return [int(t) for t in out.split()]


def off_sha256(off):
    return hashlib.sha256(np.ascontiguousarray(off, np.int64).tobytes()).hexdigest()


def partitions():
    """Vertex bins of the reference's balanced reads (-b greedy, -d linear
    partitioning) on fixture offsets; the dynamic program is O(nv^2 nparts) in
    the reference, so it runs on the smaller fixtures only."""
    cases = [("rmat", dict(scale=10, seed=1, permute=1), True),
             ("rmat", dict(scale=12, seed=3, permute=0), True),
             ("rmat", dict(scale=16, seed=1, permute=0), False),
             ("er", dict(n=2000, seed=12, p=0.01), True),
             ("rgg_lcg", dict(n=4096), True)]
    if os.path.exists(AMAZON_TSV):
        cases.append(("amazon0302", {}, False))
    out = []
    with tempfile.TemporaryDirectory() as tmp:
        for kind, prm, linear in cases:
            if kind == "rmat":
                off, _ = O.rmat_csr(prm["scale"], 16, prm["seed"], bool(prm["permute"]))
            elif kind == "er":
                off, _ = O.er_csr(prm["n"], prm["seed"], O.er_threshold(prm["p"]))
            elif kind == "rgg_lcg":
                off, _ = O.rgg_lcg(prm["n"])
            else:
                off, _ = amazon_csr()
            for nparts in (2, 3, 4, 8, 64):
                for method, name in ((0, "greedy"), (1, "linear")):
                    if method == 1 and (not linear or nparts > 8):
                        continue
                    bins = ref_partition(off, nparts, method, tmp)
                    out.append({"graph": kind, "params": prm, "nv": len(off) - 1,
                                "off_sha256": off_sha256(off), "nparts": nparts,
                                "method": name, "bins": bins})
                    print("partition", kind, prm, nparts, name, bins[:4], "...", flush=True)
    return out


def rgg_ranks():
    out = []
    with tempfile.TemporaryDirectory() as tmp:
        for n, p in ((4096, 2), (4096, 4), (4096, 8), (65536, 2), (65536, 4), (100000, 2), (100000, 4),
                     (100000, 8), (100000, 32)):
            path = os.path.join(tmp, "ranks.bin")
            subprocess.run([O.MPIEXEC, "-n", str(p), O.REF_RGG_RANKS_BIN, str(n), path], check=True,
                           capture_output=True, timeout=3600)
            a = np.fromfile(path, np.int64)
            nv, ne = int(a[0]), int(a[1])
            off, tails = a[2:3 + nv], a[3 + nv:]
            assert nv == n and len(tails) == ne and off[-1] == ne
            tri = ref_on_csr(off, tails, tmp, "ranks")
            out.append({"n": n, "nprocs": p, "seed": 1, "ne": ne, "triangles": tri,
                        "csr_sha256": csr_sha256(off, tails)})
            print("rgg ranks", n, p, ne, tri, flush=True)
    return out


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    if sys.argv[1:] == ["--ranks"]:  # refresh only the multi-rank generator fixtures
        with open(os.path.join(HERE, "golden.json")) as f:
            g = json.load(f)
        g["rgg_ranks"] = rgg_ranks()
        g["reference_rgg_ranks_binary"] = "oracle/_ref/ref_rgg_ranks"
        with open(os.path.join(HERE, "golden.json"), "w") as f:
            json.dump(g, f, indent=1)
        return
    if sys.argv[1:] == ["--partitions"]:  # refresh only the partition bins
        with open(os.path.join(HERE, "golden.json")) as f:
            g = json.load(f)
        g["partitions"] = partitions()
        g["reference_partition_binary"] = "oracle/_ref/ref_partition"
        with open(os.path.join(HERE, "golden.json"), "w") as f:
            json.dump(g, f, indent=1)
        return
    g = {"generator": "tests/golden/make_golden.py", "reference_binary": "oracle/_ref/tric_ref",
         "rgg_lcg": [], "rmat": [], "er": [], "batch": [], "amazon0302": None}
    for n in (256, 1024, 4096, 65536, 100000):
        nv, ne, tri = run_ref(["-n", str(n), "-l"])
        off, tails = O.rgg_lcg(n)
        g["rgg_lcg"].append({"n": n, "seed": 1, "ne": ne, "triangles": tri,
                             "csr_sha256": csr_sha256(off, tails)})
        print("rgg", n, ne, tri, flush=True)
    with tempfile.TemporaryDirectory() as tmp:
        for scale, seed, perm in ((8, 1, 1), (10, 1, 1), (10, 7, 0), (12, 1, 1), (12, 3, 0),
                                  (14, 1, 1), (16, 1, 1), (16, 1, 0)):
            off, tails = O.rmat_csr(scale, 16, seed, bool(perm))
            tri = ref_on_csr(off, tails, tmp, "rmat")
            g["rmat"].append({"scale": scale, "edgefactor": 16, "seed": seed, "permute": perm,
                              "abc": [0.57, 0.19, 0.19], "ne": int(off[-1]), "triangles": tri,
                              "csr_sha256": csr_sha256(off, tails)})
            print("rmat", scale, seed, perm, off[-1], tri, flush=True)
        for n, p, seed in ((300, 0.05, 11), (2000, 0.01, 12), (5000, 32.0 / 4999, 13)):
            thr = O.er_threshold(p)
            off, tails = O.er_csr(n, seed, thr)
            tri = ref_on_csr(off, tails, tmp, "er")
            g["er"].append({"n": n, "seed": seed, "thr64": thr, "ne": int(off[-1]),
                            "triangles": tri, "csr_sha256": csr_sha256(off, tails)})
            print("er", n, off[-1], tri, flush=True)
        for gi in range(8):
            n, kind, gseed, thr = O.batch_params(2026, gi)
            if kind == 0:
                off, tails = O.rgg_lcg(n, seed=(gseed ^ (gseed >> 32)) & 0xFFFFFFFF)
            else:
                off, tails = O.er_csr(n, gseed + 3, thr)
            tri = ref_on_csr(off, tails, tmp, "batch")
            g["batch"].append({"base_seed": 2026, "index": gi, "n": n, "kind": kind,
                               "ne": int(off[-1]), "triangles": tri,
                               "csr_sha256": csr_sha256(off, tails)})
            print("batch", gi, n, kind, off[-1], tri, flush=True)
        if os.path.exists(AMAZON_TSV):
            off, tails = amazon_csr()
            tri = ref_on_csr(off, tails, tmp, "amazon")
            g["amazon0302"] = {"nv": len(off) - 1, "ne": int(off[-1]), "triangles": tri,
                               "csr_sha256": csr_sha256(off, tails)}
            print("amazon", len(off) - 1, off[-1], tri, flush=True)
    g["partitions"] = partitions()
    g["reference_partition_binary"] = "oracle/_ref/ref_partition"
    g["rgg_ranks"] = rgg_ranks()
    g["reference_rgg_ranks_binary"] = "oracle/_ref/ref_rgg_ranks"
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=1)


if __name__ == "__main__":
    main()
