"""Several graphs of uneven size in one launch (pbbs_bfs_multi): every job's
parent and depth arrays equal the oracle's and a lone pbbs_bfs on its graph,
bit for bit, whatever the mix of sizes, depths, offset widths and row
orders, and also when the jobs fall back to running alone."""
import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def reversed_rows(g):
    adj = g.adj.copy()
    for v in range(g.n):
        a, b = g.offsets[v], g.offsets[v + 1]
        adj[a:b] = adj[a:b][::-1]
    return pb.HostGraph(g.offsets, adj)


def mixed():
    """(host graph, wide offsets, sources) of uneven sizes and shapes."""
    r14 = pb.rmat(14, 16 * 2 ** 14, 0.57, 0.19, 0.19, seed=3)
    r10 = pb.rmat(10, 16 * 2 ** 10, 0.57, 0.19, 0.19, seed=4)
    nz14 = np.flatnonzero(r14.degrees() > 0)
    return [
        (r14, False, [int(nz14[0]), int(nz14[100]), int(nz14[-1])]),
        (r10, True, [0, 5]),
        (pb.grid3d(20), False, [0, 777]),                       # deep: 31 levels
        (pb.grid3d(5), True, [3]),
        (pb.randlocal(5000, 4, seed=7), False, [17]),
        (pb.read_graph_from_file(f"{GOLDEN}/directed6.adj"), False, [0]),   # top-down only
        (reversed_rows(pb.rmat(11, 16 * 2 ** 11, 0.57, 0.19, 0.19, seed=8)), False, [1]),
        (pb.read_graph_from_file(f"{GOLDEN}/isolated_src.adj"), False, [0]),
    ]


def run_mix(ctx, tune=None):
    if tune:
        ctx.set_tuning(**tune)
    cases = mixed()
    dgs, srcs, hosts = [], [], []
    for g, wide, ss in cases:
        dg = ctx.upload(g, wide_offsets=wide)
        for s in ss:
            dgs.append(dg)
            srcs.append(s)
            hosts.append(g)
    say(f"{len(dgs)} jobs on {len(cases)} graphs, tuning {tune}")
    out = ctx.bfs_multi(dgs, srcs)
    for j, (p, d, st) in enumerate(out):
        wp, wd = oracle.serial_bfs(hosts[j].offsets, hosts[j].adj, srcs[j])
        assert np.array_equal(p, wp), f"job {j} (n={hosts[j].n}, source {srcs[j]}): parent"
        assert np.array_equal(d, wd), f"job {j}: depth"
        lp, ld, lst = dgs[j].bfs(srcs[j])
        assert np.array_equal(p, lp) and np.array_equal(d, ld)
        assert st["reached"] == lst["reached"] and st["edges_reached"] == lst["edges_reached"]
        assert st["levels"] == lst["levels"]
    for dg in {id(d): d for d in dgs}.values():
        dg.free()


def test_multi_mixed_graphs():
    with pb.Context(0) as ctx:
        run_mix(ctx)


@pytest.mark.parametrize("tune", [
    dict(persist_wait_us=0),                  # every job gives up at its first wait
    dict(queue_div=1_000_000),                # queues overflow into the bitmaps
    dict(td_to_bu_div=0),                     # top-down only
])
def test_multi_tuning_invariance(tune):
    with pb.Context(0) as ctx:
        run_mix(ctx, tune)


def test_multi_large_jobs_run_alone():
    """Jobs on graphs of more than 2^23 edges run alone inside the call (one
    launch sequence each on the whole device) while the small ones share one
    launch; every job still equals the oracle and a lone pbbs_bfs."""
    big = pb.rmat(19, 16 * 2 ** 19, 0.57, 0.19, 0.19, seed=6)
    assert big.m > 2 ** 23
    small = pb.rmat(12, 16 * 2 ** 12, 0.57, 0.19, 0.19, seed=2)
    nzb = np.flatnonzero(big.degrees() > 0)
    with pb.Context(0) as ctx:
        db = ctx.upload(big)
        ds = ctx.upload(small)
        dgs = [ds, db, ds, db]
        srcs = [0, int(nzb[0]), 7, int(nzb[9])]
        hosts = [small, big, small, big]
        say("large jobs alone:", [h.n for h in hosts])
        out = ctx.bfs_multi(dgs, srcs)
        for j, (p, d, st) in enumerate(out):
            wp, wd = oracle.serial_bfs(hosts[j].offsets, hosts[j].adj, srcs[j])
            assert np.array_equal(p, wp) and np.array_equal(d, wd), j
            lp, ld, _ = dgs[j].bfs(srcs[j])
            assert np.array_equal(p, lp) and np.array_equal(d, ld), j
        db.free()
        ds.free()


def test_multi_same_graph_many_sources_and_limits():
    g = pb.rmat(12, 16 * 2 ** 12, 0.57, 0.19, 0.19, seed=9)
    with pb.Context(0) as ctx:
        dg = ctx.upload(g)
        src = list(range(64))
        out = ctx.bfs_multi([dg] * 64, src, depth=False)
        for s, (p, d, _) in zip(src, out):
            wp, _ = oracle.serial_bfs(g.offsets, g.adj, s)
            assert d is None and np.array_equal(p, wp), s
        with pytest.raises(pb.PbbsError):
            ctx.bfs_multi([dg] * 65, list(range(65)))
        with pytest.raises(pb.PbbsError):
            ctx.bfs_multi([dg], [g.n])
        dg.free()
