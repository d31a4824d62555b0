"""GPU parity of the batched (multi-source) BFS: every row must equal the
oracle's single-source result for that source, bit for bit."""
import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN, golden_cases

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def check_batch(ctx, g, sources, wide=False, assume_symmetric=False, **tune):
    say(f"batch n={g.n} m={g.m} k={len(sources)} wide={wide} tune={tune}")
    dg = ctx.upload(g, wide_offsets=wide, assume_symmetric=assume_symmetric)
    try:
        P, D, st = dg.bfs_batch(sources)
    finally:
        dg.free()
    deg = g.degrees()
    tot_r = tot_m = 0
    for i, s in enumerate(sources):
        wp, wd = oracle.serial_bfs(g.offsets, g.adj, int(s))
        if not np.array_equal(P[i], wp):
            bad = np.flatnonzero(P[i] != wp)
            raise AssertionError(f"source {i}={s}: parent differs at {bad[:8]}: "
                                 f"{P[i][bad[:8]]} vs {wp[bad[:8]]}; stats {st}")
        if not np.array_equal(D[i], wd):
            bad = np.flatnonzero(D[i] != wd)
            raise AssertionError(f"source {i}={s}: depth differs at {bad[:8]}")
        tot_r += int((wp != -1).sum())
        tot_m += int(deg[wp != -1].sum())
    assert st["reached"] == tot_r
    assert st["edges_reached"] == tot_m
    return st


def test_batch_rmat_64_sources(ctx):
    g = pb.rmat(13, 16 * 2 ** 13, 0.57, 0.19, 0.19, seed=3)
    rng = np.random.default_rng(7)
    nz = np.flatnonzero(g.degrees() > 0)
    check_batch(ctx, g, rng.choice(nz, 64, replace=False))


@pytest.mark.parametrize("k", [1, 2, 17, 63])
def test_batch_sizes_and_repeats(ctx, k):
    g = pb.rmat(11, 16 * 2 ** 11, 0.57, 0.19, 0.19, seed=k)
    rng = np.random.default_rng(k)
    src = rng.integers(0, g.n, k)            # may repeat and hit isolated vertices
    if k > 1:
        src[1] = src[0]
    check_batch(ctx, g, src, wide=(k % 2 == 1))


def test_batch_torus_and_pairs(ctx):
    check_batch(ctx, pb.grid3d(12), [0, 5, 100, 1727, 999, 0])
    g = pb.read_graph_from_file(f"{GOLDEN}/pairs12.adj")
    check_batch(ctx, g, list(range(12)))


@pytest.mark.parametrize("tune", [dict(td_to_bu_div=0), dict(batch_pull_div=0),
                                  dict(batch_pull_div=1_000_000), dict(use_graphs=0),
                                  dict(grid_blocks=3), dict(grid_blocks=3, batch_pull_div=1)])
def test_batch_tuning_invariance(tune):
    g = pb.rmat(12, 16 * 2 ** 12, 0.57, 0.19, 0.19, seed=9)
    rng = np.random.default_rng(3)
    src = rng.choice(np.flatnonzero(g.degrees() > 0), 40, replace=False)
    with pb.Context(0) as c:
        c.set_tuning(**tune)
        check_batch(c, g, src)


def test_batch_rejects_bad_input(ctx):
    g = pb.rmat(8, 1000, 0.57, 0.19, 0.19, seed=2)
    dg = ctx.upload(g)
    with pytest.raises(pb.PbbsError):
        dg.bfs_batch(np.zeros(65, dtype=np.int32))
    with pytest.raises(pb.PbbsError):
        dg.bfs_batch([g.n])
    dg.free()
    d = pb.from_edges(10, [0, 1], [1, 2], symmetrise=False)
    dd = ctx.upload(d)
    with pytest.raises(pb.PbbsError, match="symmetric"):
        dd.bfs_batch([0])
    dd.free()


def shuffled_rows(g, seed):
    """Same graph with every row's entries in random order."""
    rng = np.random.default_rng(seed)
    adj = g.adj.copy()
    off = g.offsets
    for v in np.flatnonzero(np.diff(off) > 1):
        seg = adj[off[v]:off[v + 1]]
        rng.shuffle(seg)
    return pb.HostGraph(off, adj)


@pytest.mark.parametrize("tune", [dict(), dict(batch_pull_div=1_000_000), dict(grid_blocks=2)])
def test_batch_unsorted_rows(tune):
    """Rows in arbitrary order: every hit takes a max, no early exit."""
    g = shuffled_rows(pb.rmat(11, 16 * 2 ** 11, 0.57, 0.19, 0.19, seed=21), 4)
    rng = np.random.default_rng(2)
    src = rng.choice(np.flatnonzero(g.degrees() > 0), 33, replace=False)
    with pb.Context(0) as c:
        c.set_tuning(**tune)
        dg = c.upload(g, assume_symmetric=True)
        assert not dg.info["sorted"]
        dg.free()
        check_batch(c, g, src, assume_symmetric=True)


@pytest.mark.parametrize("grid", [1, 2, 5])
def test_batch_hub_rows_split_across_wavefronts(grid):
    """A star of 3000 leaves plus a path: with few wavefronts the hub's row is
    cut by slice boundaries in both pull and resolve."""
    n = 3000 + 50
    u = np.concatenate([np.zeros(2999, dtype=np.int32), np.arange(3000, 3049, dtype=np.int32),
                        [2999]])
    v = np.concatenate([np.arange(1, 3000, dtype=np.int32),
                        np.arange(3001, 3050, dtype=np.int32), [3000]])
    g = pb.from_edges(n, u, v)
    src = [5, 0, 3049, 1234, 2999, 3025, 7, 7]
    for pull in (1_000_000, 0, 10):
        with pb.Context(0) as c:
            c.set_tuning(grid_blocks=grid, batch_pull_div=pull)
            check_batch(c, g, src)


@pytest.mark.parametrize("pull", [0, 10, 1_000_000])
def test_batch_deep_levels(pull):
    """A 700-vertex path: depths reach 699, past the one-byte level table
    (levels from 255 on are written straight to the outputs)."""
    n = 700
    u = np.arange(n - 1, dtype=np.int32)
    g = pb.from_edges(n, u, u + 1)
    with pb.Context(0) as c:
        c.set_tuning(batch_pull_div=pull)
        check_batch(c, g, [0, 350, 699, 5, 0])


def test_batch_prefilled_parent_buffers():
    """Batches in a row on one context: each batch's parents come from the
    buffer filled with -1 on the side stream during the batch before, when
    it is large enough, else the clear kernel fills it.  Sizes go up and
    down, graphs change, and a batch into the caller's device buffers sits in
    between; every row equals the oracle."""
    g1 = pb.rmat(12, 16 * 4096, 0.57, 0.19, 0.19, seed=21)
    g2 = pb.grid3d(11)
    nz1 = np.flatnonzero(g1.degrees() > 0)
    with pb.Context(0) as c:
        d1, d2 = c.upload(g1), c.upload(g2)
        try:
            plan = [(d1, g1, nz1[:64]), (d1, g1, nz1[64:128]), (d1, g1, nz1[:8]),
                    (d2, g2, np.arange(0, 640, 10)), (d1, g1, nz1[200:264]),
                    ("dev", None, None), (d1, g1, nz1[300:364]), (d1, g1, nz1[:3])]
            for step, (dg, g, src) in enumerate(plan):
                if dg == "dev":
                    # caller's device buffers: the ctx buffers of a single-source run
                    # hold n entries, so use a one-source batch into them
                    s = int(nz1[5])
                    d1.bfs_device(s)
                    pp, dp = c.outputs(batched=False)
                    d1.bfs_batch_device(np.array([s], dtype=np.int32), pp, dp)
                    continue
                say(f"step {step}: n={g.n} k={len(src)}")
                P, D, _ = dg.bfs_batch(src)
                for i, s in enumerate(src):
                    wp, wd = oracle.serial_bfs(g.offsets, g.adj, int(s))
                    assert np.array_equal(P[i], wp), f"step {step} row {i}: parent"
                    assert np.array_equal(D[i], wd), f"step {step} row {i}: depth"
        finally:
            d1.free()
            d2.free()
