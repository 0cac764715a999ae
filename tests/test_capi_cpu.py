"""CPU tests of the product library that need no GPU: exports, host ingest, vite I/O.

The host ingest of libxtric (cell-grid RGG, vite reader/writer, edge-list
converter, batch mix) is checked against the oracle and the golden fixtures.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import xxtric_amd
from helpers import csr_sha256
from xxtric_amd import XtError, capi


def test_library_exports_every_declared_function():
    L = capi.lib()
    names = capi.header_functions()
    assert len(names) >= 40
    missing = [n for n in names if getattr(L, n, None) is None]
    assert missing == []
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_library_is_built_for_gfx950():
    # the embedded device code object names its target (amdgcn-amd-amdhsa--gfx950)
    blob = open(capi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_no_gpu_fails_loudly_here():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU node may be present")
    with pytest.raises(xxtric_amd.XtError):
        xxtric_amd.Context(0)


@pytest.mark.parametrize("idx", range(5))
def test_rgg_cell_grid_matches_reference_all_pairs(golden, idx):
    g = golden["rgg_lcg"][idx]
    off, tails = xxtric_amd.rgg_generate(g["n"], seed=g["seed"])
    assert off[-1] == g["ne"] and csr_sha256(off, tails) == g["csr_sha256"]


@pytest.mark.parametrize("n,seed,radius", [(3000, 7, 0.0), (5000, 3, 8.5612e-3), (2000, 11, 0.05)])
def test_rgg_cell_grid_other_seeds_and_radii(n, seed, radius):
    o1, t1 = xxtric_amd.rgg_generate(n, seed=seed, radius=radius)
    o2, t2 = O.rgg_lcg(n, seed=seed, radius=radius)
    assert (o1 == o2).all() and (t1 == t2).all()


@pytest.mark.parametrize("idx", range(9))
def test_rgg_ranks_matches_reference_generator(golden, idx):
    """xt_rgg_generate_ranks (cell grid) equals the reference's P-rank generator."""
    g = golden["rgg_ranks"][idx]
    off, tails = xxtric_amd.rgg_generate(g["n"], seed=g["seed"], nprocs=g["nprocs"])
    assert off[-1] == g["ne"] and csr_sha256(off, tails) == g["csr_sha256"]


@pytest.mark.parametrize("n,p,seed", [(6000, 2, 5), (8192, 16, 1), (3000, 8, 9)])
def test_rgg_ranks_equals_oracle_other_cases(n, p, seed):
    o1, t1 = xxtric_amd.rgg_generate(n, seed=seed, nprocs=p)
    o2, t2 = O.rgg_lcg_ranks(n, p, seed=seed)
    assert np.array_equal(o1, o2) and np.array_equal(t1, t2)


@pytest.mark.parametrize("n,p", [(4096, 3), (4098, 4), (4096, 256), (4096, 0)])
def test_rgg_ranks_rejects_what_the_reference_aborts_on(n, p):
    """graph.hpp:765-793: ranks a power of two dividing n, strips wider than the radius."""
    with pytest.raises(XtError) as e:
        xxtric_amd.rgg_generate(n, nprocs=p)
    assert e.value.code == capi.XT_ERR_ARG


def test_rgg_radius_matches_oracle():
    for n in (256, 4096, 100000, 10 ** 7):
        assert xxtric_amd.rgg_radius(n) == O.rgg_radius_ref(n)


@pytest.mark.parametrize("g", list(range(64)))
def test_batch_params_match_oracle(g):
    assert xxtric_amd.batch_params(2026, g)[:4] == O.batch_params(2026, g)


@pytest.mark.parametrize("idx", range(8))
def test_batch_graphs_match_golden(golden, idx):
    g = golden["batch"][idx]
    off, tails = xxtric_amd.batch_graph_csr(g["base_seed"], g["index"])
    assert csr_sha256(off, tails) == g["csr_sha256"]


@pytest.mark.parametrize("layout32", [False, True])
def test_vite_product_and_oracle_agree(tmp_path, layout32):
    off, tails = O.rmat_csr(10, 16, 1, True)
    p1 = str(tmp_path / "prod.bin")
    p2 = str(tmp_path / "orac.bin")
    xxtric_amd.vite_write(p1, off, tails, layout32=layout32)
    O.vite_write(p2, off, tails, layout32=layout32)
    assert open(p1, "rb").read() == open(p2, "rb").read()
    o, t = xxtric_amd.vite_read(p2, layout32=layout32)
    assert (o == off).all() and (t == tails).all()


def test_vite_reader_rebases_offsets(tmp_path):
    # graph.hpp:386-388: offsets are made relative to the first one
    p = str(tmp_path / "g.bin")
    nv, ne = 3, 4
    with open(p, "wb") as f:
        np.array([nv, ne], np.int64).tofile(f)
        np.array([10, 12, 13, 14], np.int64).tofile(f)
        rec = np.zeros(ne, dtype=[("t", np.int64), ("w", np.float64)])
        rec["t"] = [1, 2, 0, 0]
        rec.tofile(f)
    o, t = xxtric_amd.vite_read(p)
    assert list(o) == [0, 2, 3, 4] and list(t) == [1, 2, 0, 0]


TRIC = os.path.join(capi.PKG, "bin", "tric")


@pytest.mark.parametrize("layout32", [False, True])
def test_tric_generates_reference_graph(golden, tmp_path, layout32):
    g = [x for x in golden["rgg_lcg"] if x["n"] == 4096][0]
    p = str(tmp_path / "g.bin")
    args = [TRIC, "-n", "4096", "-l", "-x", "-w", p] + (["-L"] if layout32 else [])
    out = subprocess.run(args, capture_output=True, text=True, check=True).stdout
    assert "Number of vertices: 4096" in out and "Number of edges: %d" % g["ne"] in out
    off, tails = O.vite_read(p, layout32=layout32)
    assert csr_sha256(off, tails) == g["csr_sha256"]


def test_tric_rejects_bad_flags():
    r = subprocess.run([TRIC], capture_output=True, text=True)
    assert r.returncode != 0 and "Must specify some options." in r.stderr
    r = subprocess.run([TRIC, "-l"], capture_output=True, text=True)
    assert r.returncode != 0
    r = subprocess.run([TRIC, "-n", "100", "-p", "120"], capture_output=True, text=True)
    assert r.returncode != 0 and "Invalid random edge percentage" in r.stderr


def test_tric_random_extra_edges_are_seeded(tmp_path):
    outs = []
    for k in range(2):
        p = str(tmp_path / ("p%d.bin" % k))
        subprocess.run([TRIC, "-n", "2000", "-l", "-p", "5", "-x", "-w", p], capture_output=True,
                       check=True)
        outs.append(O.vite_read(p))
    (o1, t1), (o2, t2) = outs
    assert (o1 == o2).all() and (t1 == t2).all()
    base, bt = O.rgg_lcg(2000)
    assert len(t1) > len(bt)                       # edges were added
    n = len(o1) - 1                                # still simple and symmetric
    rows = np.repeat(np.arange(n), np.diff(o1))
    assert (rows != t1).all()
    fw = set(zip(rows.tolist(), t1.tolist()))
    assert all((b, a) in fw for a, b in fw)


def test_edgelist_converter(tmp_path):
    p = tmp_path / "e.tsv"
    p.write_text("# comment\n1\t2\t1\n2\t3\t1\n3\t1\t1\n3\t4\t1\n2\t1\t1\n")
    off, tails = xxtric_amd.edgelist_to_csr(str(p))
    assert list(off) == [0, 0, 2, 4, 7, 8]
    assert list(tails) == [2, 3, 1, 3, 1, 2, 4, 3]


@pytest.mark.parametrize("strict,connect", [(True, True), (True, False), (False, True)])
def test_rgg_strict_and_connect_options_match_oracle(strict, connect):
    """xt_rgg_generate_opts (cell grid) equals the oracle's all-pairs loop with the
    same test (< or <=) and radius (sqrt(2 ln n / (pi n)) or graph.hpp:789-791)."""
    n = 6000
    r = xxtric_amd.rgg_radius_connect(n) if connect else 0.0
    o1, t1 = xxtric_amd.rgg_generate(n, seed=9, strict=strict, connect=connect)
    o2, t2 = O.rgg_lcg(n, seed=9, radius=r, strict=strict)
    assert (o1 == o2).all() and (t1 == t2).all()


def test_rgg_strict_drops_a_pair_at_exactly_the_radius():
    n = 3000
    X, Y = O.rgg_lcg_coords(n, 4)
    dx, dy = X[7] - X, Y[7] - Y
    d = np.sqrt(dx * dx + dy * dy)
    j = int(np.argsort(d)[5])
    for strict in (False, True):
        o, t = xxtric_amd.rgg_generate(n, seed=4, radius=float(d[j]), strict=strict)
        assert (j in set(t[o[7]:o[8]].tolist())) == (not strict)
