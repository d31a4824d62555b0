"""PBBS text formats (AdjacencyGraph, sequenceInt), CPU only."""
import numpy as np
import pytest

import pbbsbench_amd as pb
from conftest import GOLDEN


def test_adjacency_graph_round_trip(tmp_path):
    g = pb.rmat(9, 3000, 0.57, 0.19, 0.19, seed=3)
    p = tmp_path / "g.adj"
    pb.write_graph_to_file(g, p)
    text = p.read_text().split("\n")
    assert text[0] == "AdjacencyGraph"
    assert int(text[1]) == g.n and int(text[2]) == g.m
    h = pb.read_graph_from_file(p)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.adj, h.adj)


def test_reader_matches_reference_layout(tmp_path):
    # Tokens may be separated by any run of space, tab, LF, CR (common/IO.h:41-50).
    p = tmp_path / "t.adj"
    p.write_bytes(b"AdjacencyGraph\r\n3 \t4\n0\n1 3\n\n1\t0 2\r\n1\n")
    g = pb.read_graph_from_file(p)
    assert list(g.offsets) == [0, 1, 3, 4]
    assert list(g.adj) == [1, 0, 2, 1]


@pytest.mark.parametrize("body,match", [
    (b"EdgeArray\n1\n0\n0\n", "header"),
    (b"AdjacencyGraph\n2\n1\n0\n1\n", "n\\+m\\+3"),           # one token short
    (b"AdjacencyGraph\n2\n1\n0\n1\n1\n5\n", "n\\+m\\+3"),     # one token long
    (b"AdjacencyGraph\n2\n1\n0\nx\n1\n", "integer"),
])
def test_reader_rejects_malformed(tmp_path, body, match):
    p = tmp_path / "bad.adj"
    p.write_bytes(body)
    with pytest.raises(pb.PbbsError, match=match):
        pb.read_graph_from_file(p)


def test_reader_missing_file(tmp_path):
    with pytest.raises(pb.PbbsError, match="cannot open"):
        pb.read_graph_from_file(tmp_path / "nope.adj")


def test_sequence_int_round_trip(tmp_path):
    a = np.array([-1, 0, 5, 2147483647, -2147483648, 17], dtype=np.int32)
    p = tmp_path / "s.txt"
    pb.write_sequence_to_file(a, p)
    assert p.read_text() == "sequenceInt\n-1\n0\n5\n2147483647\n-2147483648\n17\n"
    assert np.array_equal(pb.read_int_seq_from_file(p), a)
    big = np.random.default_rng(1).integers(-1, 10 ** 7, 3_000_000).astype(np.int32)
    pb.write_sequence_to_file(big, p)
    assert np.array_equal(pb.read_int_seq_from_file(p), big)


def test_sequence_reader_header(tmp_path):
    p = tmp_path / "s.txt"
    p.write_text("sequenceDouble\n1.0\n")
    with pytest.raises(pb.PbbsError, match="sequenceInt"):
        pb.read_int_seq_from_file(p)


def test_golden_files_parse():
    g = pb.read_graph_from_file(f"{GOLDEN}/torus4.adj")
    assert g.n == 64 and g.m == 384
    assert pb.read_int_seq_from_file(f"{GOLDEN}/torus4.parent").size == 64


def test_reader_random_separators_and_edges(tmp_path):
    """Runs of mixed separators, no trailing newline, tokens falling on the
    parser's chunk cuts: the parse must equal the tokens written."""
    rng = np.random.default_rng(7)
    g = pb.rmat(12, 40000, 0.57, 0.19, 0.19, seed=2)
    toks = ["AdjacencyGraph", str(g.n), str(g.m)] + [str(x) for x in g.offsets[:-1]] + \
        [str(x) for x in g.adj]
    seps = [" ", "\t", "\n", "\r", "\r\n", "  \t"]
    pick = rng.integers(0, len(seps), len(toks))
    body = "".join(t + seps[k] for t, k in zip(toks, pick)).rstrip()
    p = tmp_path / "r.adj"
    p.write_text("\n \t" + body)
    h = pb.read_graph_from_file(p)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.adj, h.adj)


@pytest.mark.parametrize("body,n,m", [
    (b"AdjacencyGraph\n1\n0\n0", 1, 0),
    (b"AdjacencyGraph 1 0 0", 1, 0),
    (b"AdjacencyGraph\n2\n2\n0\n1\n1\n0\n", 2, 2),
])
def test_reader_small_files(tmp_path, body, n, m):
    p = tmp_path / "s.adj"
    p.write_bytes(body)
    g = pb.read_graph_from_file(p)
    assert g.n == n and g.m == m and g.offsets[-1] == m


@pytest.mark.parametrize("body,match", [
    (b"AdjacencyGraph\n2\n1\n0\n1\n12abc\n", "integer"),
    (b"AdjacencyGraph\n2\n1\n0\n1\n1-2\n", "integer"),
    (b"AdjacencyGraph\n2\n1\n0\n1\n9999999999\n", "integer"),       # beyond int32
    (b"AdjacencyGraphX\n1\n0\n0\n", "header"),
    (b"AdjacencyGraph\n", "missing n and m"),
])
def test_reader_rejects_bad_tokens(tmp_path, body, match):
    p = tmp_path / "bad.adj"
    p.write_bytes(body)
    with pytest.raises(pb.PbbsError, match=match):
        pb.read_graph_from_file(p)


def test_binary_round_trip(tmp_path):
    """text -> host CSR -> binary -> host CSR, on a generated graph and on
    every golden graph."""
    import glob
    import os
    graphs = [pb.rmat(13, 80000, 0.5, 0.1, 0.1, seed=4, permute_seed=9)]
    graphs += [pb.read_graph_from_file(f) for f in sorted(glob.glob(os.path.join(GOLDEN, "*.adj")))]
    for i, g in enumerate(graphs):
        p = tmp_path / f"g{i}.csr"
        pb.write_graph_binary(g, p)
        raw = p.read_bytes()
        assert raw[:8] == b"PBBSCSR1" and len(raw) == 32 + 8 * (g.n + 1) + 4 * g.m
        assert int.from_bytes(raw[8:16], "little") == g.n
        assert int.from_bytes(raw[16:24], "little") == g.m
        assert np.array_equal(np.frombuffer(raw[32:32 + 8 * (g.n + 1)], "<i8"), g.offsets)
        h = pb.read_graph_binary(p)
        assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.adj, h.adj)


def test_binary_rejects_bad_files(tmp_path):
    g = pb.grid3d(3)
    p = tmp_path / "g.csr"
    pb.write_graph_binary(g, p)
    raw = p.read_bytes()
    (tmp_path / "short.csr").write_bytes(raw[:-4])
    with pytest.raises(pb.PbbsError, match="bytes"):
        pb.read_graph_binary(tmp_path / "short.csr")
    (tmp_path / "magic.csr").write_bytes(b"PBBSCSR2" + raw[8:])
    with pytest.raises(pb.PbbsError, match="not a PBBSCSR1"):
        pb.read_graph_binary(tmp_path / "magic.csr")
    (tmp_path / "text.adj").write_text("AdjacencyGraph\n1\n0\n0\n")
    with pytest.raises(pb.PbbsError, match="not a PBBSCSR1"):
        pb.read_graph_binary(tmp_path / "text.adj")


def test_edge_array_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    u = rng.integers(0, 1000, 5000).astype(np.int32)
    v = rng.integers(0, 1000, 5000).astype(np.int32)
    p = tmp_path / "e"
    pb.write_edge_array_to_file(u, v, p)
    lines = open(p).read().split("\n")
    assert lines[0] == "EdgeArray" and lines[1] == f"{u[0]} {v[0]}"
    n, u2, v2 = pb.read_edge_array_from_file(p)
    assert n == max(u.max(), v.max()) + 1
    assert np.array_equal(u2, u) and np.array_equal(v2, v)


@pytest.mark.parametrize("body,n,m", [
    ("EdgeArray\n", 1, 0),                      # no edges: n = 0 + 1, as the reference's reduce
    ("EdgeArray 3 4\t5\r\n6", 7, 2),            # any separators
    ("EdgeArray\n1 2\n3", 3, 1),                # an odd last token is ignored
])
def test_edge_array_reader_small_files(tmp_path, body, n, m):
    p = tmp_path / "e"
    p.write_text(body)
    got_n, u, v = pb.read_edge_array_from_file(p)
    assert (got_n, u.size, v.size) == (n, m, m)


@pytest.mark.parametrize("body,match", [
    ("AdjacencyGraph\n1 2\n", "missing header"),
    ("EdgeArray\n1 x\n", "edge 0"),
    ("EdgeArray\n0 1\n-3 2\n", "edge 1"),
    ("EdgeArray\n0 2147483647\n", "edge 0"),
])
def test_edge_array_reader_rejects_malformed(tmp_path, body, match):
    p = tmp_path / "e"
    p.write_text(body)
    with pytest.raises(pb.PbbsError, match=match):
        pb.read_edge_array_from_file(p)
