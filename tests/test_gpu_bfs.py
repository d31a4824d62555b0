"""GPU parity: the HIP BFS (through the C ABI) against the oracle, bit-exact.

Every case prints its name before touching the GPU so a failing run names it.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN, golden_cases

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def run_and_compare(ctx, g, source, wide=False, assume_symmetric=False, want=None):
    say(f"n={g.n} m={g.m} src={source} wide={wide}")
    dg = ctx.upload(g, wide_offsets=wide, assume_symmetric=assume_symmetric)
    try:
        parent, depth, st = dg.bfs(source)
    finally:
        dg.free()
    if want is None:
        want = oracle.serial_bfs(g.offsets, g.adj, source)
    wp, wd = want
    if not np.array_equal(parent, wp):
        bad = np.flatnonzero(parent != wp)
        raise AssertionError(f"parent differs at {bad[:8]} (got {parent[bad[:8]]}, "
                             f"want {wp[bad[:8]]}); stats {st}")
    if not np.array_equal(depth, wd):
        bad = np.flatnonzero(depth != wd)
        raise AssertionError(f"depth differs at {bad[:8]} (got {depth[bad[:8]]}, "
                             f"want {wd[bad[:8]]}); stats {st}")
    assert st["reached"] == int((wp != -1).sum())
    deg = g.degrees()
    assert st["edges_reached"] == int(deg[wp != -1].sum())
    assert st["levels"] == int(wd[wd < g.n].max()) + 1
    return st


@pytest.mark.parametrize("name,meta", golden_cases(), ids=[c[0] for c in golden_cases()])
@pytest.mark.parametrize("wide", [False, True])
def test_golden(ctx, name, meta, wide):
    g = pb.read_graph_from_file(f"{GOLDEN}/{name}.adj")
    wp = pb.read_int_seq_from_file(f"{GOLDEN}/{name}.parent")
    wd = pb.read_int_seq_from_file(f"{GOLDEN}/{name}.depth")
    run_and_compare(ctx, g, meta["source"], wide=wide, want=(wp, wd))


def test_rmat16_config1(ctx):
    # RMAT 2^16 vertices, 2^19 draws, (a,b,c) = (.5,.1,.1), source 0.
    g = pb.rmat(16, 2 ** 19, 0.5, 0.1, 0.1, seed=1)
    st = run_and_compare(ctx, g, 0, wide=False)
    p, _ = oracle.det_bfs(g.offsets, g.adj, 0)
    assert oracle.check_bfs(g.offsets, g.adj, 0, p) == -1


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("wide", [False, True])
def test_rmat_skewed_random_sources(ctx, seed, wide):
    g = pb.rmat(14, 16 * 2 ** 14, 0.57, 0.19, 0.19, seed=seed)
    rng = np.random.default_rng(seed)
    nz = np.flatnonzero(g.degrees() > 0)
    for s in rng.choice(nz, 3, replace=False):
        run_and_compare(ctx, g, int(s), wide=wide)


@pytest.mark.parametrize("d", [3, 5, 17, 40])
def test_torus_closed_form(ctx, d):
    g = pb.grid3d(d)
    run_and_compare(ctx, g, 0, want=oracle.torus_closed_form(d))


def test_randlocal(ctx):
    g = pb.randlocal(200_000, 10, seed=3)
    run_and_compare(ctx, g, 0, wide=True)
    run_and_compare(ctx, g, 123_457, wide=False)


def test_permuted_rmat(ctx):
    g = pb.rmat(13, 16 * 2 ** 13, 0.57, 0.19, 0.19, seed=9, permute_seed=5)
    nz = np.flatnonzero(g.degrees() > 0)
    run_and_compare(ctx, g, int(nz[len(nz) // 2]))


@pytest.mark.parametrize("tune", [
    dict(td_to_bu_div=0),                       # top-down only
    dict(td_to_bu_div=1_000_000, bu_to_td_div=1_000_000),  # bottom-up as early as possible
    dict(td_to_bu_div=1_000_000, bu_to_td_div=1),  # bottom-up once entered
    dict(bu_serial_depth=0),
    dict(bu_serial_depth=16),
    dict(bu_serial_depth=4096),
    dict(grid_blocks=1),
    dict(grid_blocks=7),
    dict(use_graphs=0),
    dict(chunk_levels=1),
    dict(bu_row_tops=0),                        # bottom-up reads row tops from the CSR
    dict(bu_row_tops=0, td_to_bu_div=1_000_000, bu_to_td_div=1_000_000),
    dict(bu_row_tops=0, bu_serial_depth=0, grid_blocks=3),
    # frontier queues of 64 entries per sub-queue: larger frontiers overflow and
    # the level after them expands the frontier bitmap
    dict(queue_div=1_000_000),
    dict(queue_div=1_000_000, td_to_bu_div=0),
    dict(queue_div=1_000_000, td_to_bu_div=0, grid_blocks=7, use_graphs=0),
    # every level in one persistent launch (grid barrier between levels), on
    # grids of 1, 3 and 7 workgroups and the default; and a persistent launch
    # that gives up at its first wait (the BFS then runs again one level per
    # launch)
    dict(persist=1),
    dict(persist=1, grid_blocks=1),
    dict(persist=1, grid_blocks=3, queue_div=1_000_000),
    dict(persist=1, grid_blocks=7, td_to_bu_div=0),
    dict(persist=1, persist_wait_us=0),
    dict(persist=1, persist_wait_us=0, grid_blocks=7),
])
def test_tuning_invariance(tune):
    say("tuning", tune)
    g = pb.rmat(13, 16 * 2 ** 13, 0.57, 0.19, 0.19, seed=11)
    want = oracle.serial_bfs(g.offsets, g.adj, 1)
    with pb.Context(0) as c:
        c.set_tuning(**tune)
        run_and_compare(c, g, 1, want=want)
        t = pb.grid3d(9)
        run_and_compare(c, t, 0, want=oracle.torus_closed_form(9))


def test_directed_graph_topdown_only(ctx):
    rng = np.random.default_rng(5)
    n = 3000
    u = rng.integers(0, n, 20000)
    v = rng.integers(0, n, 20000)
    g = pb.from_edges(n, u, v, symmetrise=False)
    dg = ctx.upload(g)
    assert not dg.info["symmetric"]
    dg.free()
    run_and_compare(ctx, g, int(u[0]))


def test_unsorted_rows_full_scan(ctx):
    # Same symmetric graph with every row reversed: bottom-up must scan whole
    # rows (no early exit) and still pick the largest frontier neighbour.
    g = pb.rmat(12, 16 * 2 ** 12, 0.57, 0.19, 0.19, seed=4)
    adj = g.adj.copy()
    for v in range(g.n):
        a, b = g.offsets[v], g.offsets[v + 1]
        adj[a:b] = adj[a:b][::-1]
    h = pb.HostGraph(g.offsets, adj)
    dg = ctx.upload(h)
    info = dg.info
    dg.free()
    assert not info["sorted"]
    want = oracle.serial_bfs(g.offsets, g.adj, 0)
    run_and_compare(ctx, h, 0, assume_symmetric=True, want=want)


def test_edge_shapes(ctx):
    # single vertex, no edges
    run_and_compare(ctx, pb.HostGraph(np.array([0, 0]), np.zeros(0, np.int32)), 0)
    # n not a multiple of 64, path graph (many levels)
    n = 1000
    g = pb.from_edges(n, np.arange(n - 1), np.arange(1, n))
    run_and_compare(ctx, g, 0)
    run_and_compare(ctx, g, n - 1)
    # star with a hub of degree 50k: one row split over many wavefronts
    n = 50_001
    g = pb.from_edges(n, np.zeros(n - 1, np.int64), np.arange(1, n))
    run_and_compare(ctx, g, 0)
    run_and_compare(ctx, g, 777)


def test_persistent_rebuilds_bitmaps():
    """A persistent launch keeps no frontier or visited bitmap while its levels
    are top-down with complete queues, and rebuilds them from the depths at
    the first level that needs them: level 0 (the hub of a star turns
    bottom-up at once), level 1 (from a leaf of the star), level 200 (a path
    of 200 vertices into an RMAT graph), and a level whose queue overflowed
    (queues of 64 entries per sub-queue)."""
    star_n = 50_001
    star = pb.from_edges(star_n, np.zeros(star_n - 1, np.int64), np.arange(1, star_n))
    r = pb.rmat(13, 16 * 2 ** 13, 0.57, 0.19, 0.19, seed=11)
    lp = 200
    ru = np.repeat(np.arange(r.n, dtype=np.int64), r.degrees()) + lp
    rv = np.asarray(r.adj, dtype=np.int64) + lp
    pu = np.concatenate([np.arange(lp - 1), [lp - 1], ru])
    pv = np.concatenate([np.arange(1, lp), [lp], rv])
    path_rmat = pb.from_edges(lp + r.n, pu, pv)
    for tune, g, s, min_bu in [
        (dict(persist=1), star, 0, 1),
        (dict(persist=1), star, 777, 1),
        (dict(persist=1), path_rmat, 0, 1),
        (dict(persist=1, grid_blocks=7), path_rmat, 0, 1),
        (dict(persist=1, queue_div=1_000_000, td_to_bu_div=0), path_rmat, 0, 0),
    ]:
        say("persistent bitmaps", tune, g.n, s)
        with pb.Context(0) as c:
            c.set_tuning(**tune)
            st = run_and_compare(c, g, s)
        assert st["launches"] == 1 and st["bu_levels"] >= min_bu, st


def test_repeat_determinism(ctx):
    g = pb.rmat(15, 16 * 2 ** 15, 0.57, 0.19, 0.19, seed=21)
    dg = ctx.upload(g)
    say("repeat determinism")
    ref = dg.bfs(0)
    for _ in range(5):
        p, d, _ = dg.bfs(0)
        assert np.array_equal(p, ref[0]) and np.array_equal(d, ref[1])
    dg.free()


def test_invalid_inputs(ctx):
    from pbbsbench_amd import _lib
    g = pb.rmat(8, 1000, 0.57, 0.19, 0.19, seed=2)
    dg = ctx.upload(g)
    with pytest.raises(pb.PbbsError):
        dg.bfs(-1)
    with pytest.raises(pb.PbbsError):
        dg.bfs(g.n)
    # unknown flag bits are rejected by every BFS entry point
    lib = _lib.lib()
    assert lib.pbbs_bfs(ctx._h, dg._h, 0, None, None, 0x80, None) == _lib.PBBS_ERR_INVALID
    src = (C.c_int32 * 1)(0)
    assert lib.pbbs_bfs_batch(ctx._h, dg._h, src, 1, None, None, 0x80, None) == \
        _lib.PBBS_ERR_INVALID
    dg.free()
    bad = pb.HostGraph(g.offsets, g.adj.copy())
    bad.adj[5] = g.n + 3
    with pytest.raises(pb.PbbsError, match="outside"):
        ctx.upload(bad)
    off = g.offsets.copy()
    off[10] = off[11] + 5   # decreasing at 10 -> 11
    with pytest.raises(pb.PbbsError, match="offsets"):
        ctx.upload(pb.HostGraph(off, g.adj))


def test_graph_outlives_its_context():
    """A graph whose uploading context is destroyed stays usable from another
    context and can still be downloaded and freed."""
    g = pb.rmat(10, 8 * 2 ** 10, 0.57, 0.19, 0.19, seed=3)
    want = oracle.serial_bfs(g.offsets, g.adj, 0)
    c1 = pb.Context(0)
    dg = c1.upload(g)
    c1.close()
    with pb.Context(0) as c2:
        dg.ctx = c2
        parent, depth, _ = dg.bfs(0)
        assert np.array_equal(parent, want[0]) and np.array_equal(depth, want[1])
        d = dg.download()
        assert np.array_equal(d.adj, g.adj)
    dg.free()


def test_first_chunk_shorter_than_the_bfs(ctx):
    """The first level chunk is sized from the graph's previous run: a deeper
    BFS after a shallow one must continue with further chunks, and a shallow
    one after a deep one must stop at its own end."""
    n = 40
    u = np.arange(n - 1)
    g = pb.from_edges(n, u, u + 1)                 # path 0 - 1 - ... - 39
    dg = ctx.upload(g)
    try:
        for s in (20, 0, 39, 20, 3):
            say("path source", s)
            parent, depth, st = dg.bfs(s)
            wp, wd = oracle.serial_bfs(g.offsets, g.adj, s)
            assert np.array_equal(parent, wp) and np.array_equal(depth, wd)
            assert st["levels"] == int(wd.max()) + 1
    finally:
        dg.free()
