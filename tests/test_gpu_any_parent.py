"""Any-parent BFS (PBBS_BFS_ANY_PARENT): the contract of
backForwardBFS/BFS.C:38-61, whose compare-and-swap from -1 lets whichever
frontier neighbour claims a vertex first become its parent.

Checks, on the same cases as the deterministic path:
- depths equal the oracle's (serial BFS levels);
- every reached vertex other than the source has a parent u with (u, v) an
  edge and depth[u] = depth[v] - 1; the source is its own parent; unreached
  vertices have -1 (this is bench/BFSCheck.C:59-90, also run through the
  oracle's restatement);
- repeated runs give the same outputs;
- top-down levels keep the largest parent, so a BFS with no bottom-up level
  equals the deterministic one bit for bit."""
import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN, golden_cases

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def tree_ok(g, source, parent, depth):
    """Every reached v != source: parent p with p -> v an edge, depth[p] = depth[v] - 1."""
    off = np.asarray(g.offsets, dtype=np.int64)
    adj = np.asarray(g.adj, dtype=np.int64)
    n = off.size - 1
    assert parent[source] == source and depth[source] == 0
    reached = np.flatnonzero(parent != -1)
    assert np.array_equal(reached, np.flatnonzero(depth < n))
    v = reached[reached != source]
    p = parent[v].astype(np.int64)
    assert (p >= 0).all() and (p < n).all()
    assert np.array_equal(depth[p], depth[v] - 1)
    # (p, v) is an edge: look the key p*n + v up among the sorted edge keys
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    keys = np.sort(src * n + adj)
    want = p * n + v
    pos = np.searchsorted(keys, want)
    hit = (pos < keys.size) & (keys[np.minimum(pos, keys.size - 1)] == want)
    assert hit.all(), f"no edge parent -> v at {v[~hit][:8]}"


def run_any(ctx, g, source, wide=False, assume_symmetric=False, min_bu=0):
    say(f"any-parent n={g.n} m={g.m} src={source} wide={wide}")
    dg = ctx.upload(g, wide_offsets=wide, assume_symmetric=assume_symmetric)
    symmetric = dg.info["symmetric"]
    try:
        parent, depth, st = dg.bfs(source, any_parent=True)
        p2, d2, _ = dg.bfs(source, any_parent=True)
        pd, dd, std = dg.bfs(source)
    finally:
        dg.free()
    _, wd = oracle.serial_bfs(g.offsets, g.adj, source)
    assert np.array_equal(depth, wd)
    assert np.array_equal(p2, parent) and np.array_equal(d2, depth)
    if symmetric:   # BFSCheck reads N(v) as v's in-neighbours
        assert oracle.check_bfs(g.offsets, g.adj, source, parent) == -1
    tree_ok(g, source, parent, depth)
    assert st["bu_levels"] >= min_bu, st
    if st["bu_levels"] == 0 and std["bu_levels"] == 0:
        assert np.array_equal(parent, pd)
    return parent, pd, st


@pytest.mark.parametrize("name,meta", golden_cases(), ids=[c[0] for c in golden_cases()])
def test_any_parent_golden(ctx, name, meta):
    g = pb.read_graph_from_file(f"{GOLDEN}/{name}.adj")
    run_any(ctx, g, meta["source"])


@pytest.mark.parametrize("wide", [False, True])
def test_any_parent_rmat(ctx, wide):
    g = pb.rmat(14, 16 * 2 ** 14, 0.57, 0.19, 0.19, seed=5)
    src = int(np.flatnonzero(g.degrees() > 0)[3])
    parent, det, st = run_any(ctx, g, src, wide=wide, min_bu=1)
    # bottom-up levels took the smallest frontier neighbour here, not the largest
    assert not np.array_equal(parent, det)


def test_any_parent_unsorted_rows(ctx):
    """Rows reversed (unsorted): the probe from the row start still stops at the
    first frontier neighbour, which is a valid parent whatever the order."""
    g = pb.rmat(12, 16 * 2 ** 12, 0.57, 0.19, 0.19, seed=9)
    adj = g.adj.copy()
    for v in range(g.n):
        adj[g.offsets[v]:g.offsets[v + 1]] = adj[g.offsets[v]:g.offsets[v + 1]][::-1]
    h = pb.HostGraph(g.offsets, adj)
    run_any(ctx, h, int(np.flatnonzero(g.degrees() > 0)[0]), assume_symmetric=True, min_bu=1)


def test_any_parent_shapes(ctx):
    # torus: a valid tree with the closed form's depths
    t = pb.grid3d(9)
    run_any(ctx, t, 0)
    # star: the hub turns bottom-up at once; from a leaf, one level later
    n = 50_001
    star = pb.from_edges(n, np.zeros(n - 1, np.int64), np.arange(1, n))
    run_any(ctx, star, 0, min_bu=1)
    run_any(ctx, star, 777, min_bu=1)
    # directed graph: top-down only
    rng = np.random.default_rng(5)
    u = rng.integers(0, 3000, 20000)
    v = rng.integers(0, 3000, 20000)
    run_any(ctx, pb.from_edges(3000, u, v, symmetrise=False), int(u[0]))


@pytest.mark.parametrize("tune", [
    dict(persist=1),
    dict(persist=1, grid_blocks=3),
    dict(persist=1, persist_wait_us=0),         # gives up at its first wait; runs again per level
    dict(grid_blocks=7, use_graphs=0),
    dict(bu_serial_depth=0),                    # parked rows scanned by the whole wavefront
    dict(queue_div=1_000_000),
    dict(td_to_bu_div=1_000_000, bu_to_td_div=1_000_000),   # bottom-up from level 0 on
])
def test_any_parent_tuning(tune):
    g = pb.rmat(13, 16 * 2 ** 13, 0.57, 0.19, 0.19, seed=11)
    with pb.Context(0) as c:
        c.set_tuning(**tune)
        say("any-parent tuning", tune)
        run_any(c, g, 1)
        run_any(c, pb.grid3d(9), 0)


def test_any_parent_full_size():
    """rmat24 from two sources: the GPU checker finds a valid tree with exact
    depths (bench/BFSCheck.C:59-90), and the depths equal the deterministic run's."""
    with pb.Context(0) as c:
        say("any-parent full size: rmat24")
        dg = c.gen_rmat(24, 2 ** 28, 0.57, 0.19, 0.19, seed=1, wide_offsets=True)
        try:
            deg = dg.download().degrees()
            for s in np.flatnonzero(deg > 0)[[5, 1000]]:
                p, d, st = dg.bfs(int(s), any_parent=True)
                r = dg.check(int(s), p, d)
                assert r["valid"] == 1 and r["depth_bad"] == 0 and r["bad"] == 0, r
                pd, dd, _ = dg.bfs(int(s))
                assert np.array_equal(d, dd)
                assert st["bu_levels"] >= 1
        finally:
            dg.free()
