"""pbbs_graph_read: an AdjacencyGraph text file or a binary CSR file read
straight into HBM must hold the same CSR as the host reader's output
uploaded, and BFS on it must equal the oracle."""
import os

import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("fmt", ["text", "binary"])
@pytest.mark.parametrize("wide", [False, True])
def test_read_equals_host(ctx, tmp_path, fmt, wide):
    g = pb.rmat(14, 16 * 2 ** 14, 0.57, 0.19, 0.19, seed=8, permute_seed=3)
    p = tmp_path / ("g.adj" if fmt == "text" else "g.csr")
    (pb.write_graph_to_file if fmt == "text" else pb.write_graph_binary)(g, p)
    say(fmt, "wide" if wide else "narrow", g.n, g.m)
    dg = ctx.read_graph(p, wide_offsets=wide)
    try:
        d = dg.download()
        assert d.n == g.n and d.m == g.m
        assert np.array_equal(d.offsets, g.offsets) and np.array_equal(d.adj, g.adj)
        info = dg.info
        assert info["offsets_i64"] == int(wide) and info["symmetric"] == 1
        src = int(np.flatnonzero(g.degrees() > 0)[5])
        parent, depth, _ = dg.bfs(src)
        wp, wd = oracle.serial_bfs(g.offsets, g.adj, src)
        assert np.array_equal(parent, wp) and np.array_equal(depth, wd)
    finally:
        dg.free()


def test_read_golden_graphs(ctx, tmp_path):
    """Every golden graph, read from its text file and from a binary copy,
    gives the golden parents."""
    for name in ("bintree31", "layered11", "torus4", "comb10", "pairs12"):
        path = os.path.join(GOLDEN, f"{name}.adj")
        wp = pb.read_int_seq_from_file(os.path.join(GOLDEN, f"{name}.parent"))
        s = int(np.flatnonzero(wp == np.arange(wp.size))[0])
        h = pb.read_graph_from_file(path)
        bpath = tmp_path / f"{name}.csr"
        pb.write_graph_binary(h, bpath)
        for p in (path, bpath):
            say(name, p)
            dg = ctx.read_graph(p)
            try:
                parent, _, _ = dg.bfs(s)
                assert np.array_equal(parent, wp)
            finally:
                dg.free()


def test_read_rejects_bad_files(ctx, tmp_path):
    bad = tmp_path / "bad.adj"
    bad.write_text("AdjacencyGraph\n2\n2\n0\n4294967297\n1\n0\n")   # offset wraps in 32 bits
    with pytest.raises(pb.PbbsError, match="outside"):
        ctx.read_graph(bad)
    bad.write_text("AdjacencyGraph\n2\n1\n0\n1\n5\n")                 # neighbour id >= n
    with pytest.raises(pb.PbbsError, match="outside"):
        ctx.read_graph(bad)
    with pytest.raises(pb.PbbsError, match="cannot open"):
        ctx.read_graph(tmp_path / "missing.adj")
