"""GPU tests of the serving loop (pbbs_server_*): every query's parent and
depth arrays must equal the oracle's lone single-source BFS bit for bit,
however the server batched, overlapped or queued it."""
import threading

import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def expect(g, s, cache):
    key = (id(g), int(s))
    if key not in cache:
        cache[key] = oracle.serial_bfs(g.offsets, g.adj, int(s))
    return cache[key]


def check_results(graphs, got, queries, cache):
    batched = 0
    for t, (gid, s) in queries.items():
        r, p, d = got[t]
        assert r["status"] == 0, r
        assert r["graph"] == gid and r["source"] == s
        assert r["latency_ms"] >= r["queue_ms"] >= 0
        wp, wd = expect(graphs[gid], s, cache)
        if not np.array_equal(p, wp):
            bad = np.flatnonzero(p != wp)
            raise AssertionError(f"graph {gid} source {s} (batch {r['batch']}): parent differs "
                                 f"at {bad[:8]}: {p[bad[:8]]} vs {wp[bad[:8]]}")
        if d is not None:
            assert np.array_equal(d, wd), f"graph {gid} source {s}: depth differs"
        batched += r["batch"] > 1
    return batched


def mixed_graphs():
    return [
        pb.rmat(12, 16 * 2 ** 12, 0.57, 0.19, 0.19, seed=5),
        pb.grid3d(10),
        pb.read_graph_from_file(f"{GOLDEN}/directed6.adj"),   # not symmetric: never merged
        pb.randlocal(3000, 5, seed=2),
    ]


def test_server_mix_equals_alone():
    """Queries on four graphs submitted in one burst; the symmetric graphs'
    queries get merged into batched runs, the directed one's run alone."""
    graphs = mixed_graphs()
    rng = np.random.default_rng(11)
    cache = {}
    with pb.Server(0, slots=3) as srv:
        ids = [srv.add_graph(g) for g in graphs]
        queries = {}
        for i in range(150):
            gid = ids[int(rng.integers(0, len(ids)))]
            s = int(rng.integers(0, graphs[gid].n))
            say(f"submit {i}: graph {gid} source {s}")
            queries[srv.submit(gid, s, depth=(i % 3 != 0))] = (gid, s)
        got = srv.drain(queries)
        assert srv.load() == {"queued": 0, "running": 0, "hbm_in_use": 0}
    batched = check_results(graphs, got, queries, cache)
    assert batched > 0, "a burst of queries on one graph should be merged"
    for t, (gid, _) in queries.items():
        if gid == ids[2]:
            assert got[t][0]["batch"] == 1


def test_server_alone_then_mixed_same_answers():
    """The same query answered alone (max_batch 1, one slot) and inside a
    concurrent mix gives identical arrays."""
    g = pb.rmat(11, 16 * 2 ** 11, 0.57, 0.19, 0.19, seed=8)
    src = [0, 17, 17, 400, 2047, 5]
    alone = {}
    with pb.Server(0, slots=1, max_batch=1) as srv:
        gid = srv.add_graph(g, wide_offsets=True)
        for s in src:
            t = srv.submit(gid, s)
            r, p, d = srv.drain([t])[t]
            assert r["batch"] == 1 and r["status"] == 0
            alone[s] = (p, d)
    with pb.Server(0, slots=2, max_batch=4) as srv:
        gid = srv.add_graph(g)
        other = srv.add_graph(pb.grid3d(8))
        ts = {}
        for rep in range(3):
            for s in src:
                ts[srv.submit(gid, s)] = s
                srv.submit(other, (s * 7) % 512)
        got = srv.drain(ts)
    for t, s in ts.items():
        r, p, d = got[t]
        assert r["status"] == 0 and r["batch"] <= 4
        assert np.array_equal(p, alone[s][0]) and np.array_equal(d, alone[s][1]), s
    wp, wd = oracle.serial_bfs(g.offsets, g.adj, 17)
    assert np.array_equal(alone[17][0], wp) and np.array_equal(alone[17][1], wd)


def test_server_concurrent_submitters():
    """Four host threads submit while the worker runs."""
    graphs = mixed_graphs()[:2]
    cache = {}
    with pb.Server(0, slots=4) as srv:
        ids = [srv.add_graph(g) for g in graphs]
        queries = {}
        lock = threading.Lock()

        def client(seed):
            rng = np.random.default_rng(seed)
            for _ in range(25):
                gid = ids[int(rng.integers(0, 2))]
                s = int(rng.integers(0, graphs[gid].n))
                t = srv.submit(gid, s)
                with lock:
                    queries[t] = (gid, s)

        th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        say(f"{len(queries)} queries from 4 threads")
        got = srv.drain(queries)
    check_results(graphs, got, queries, cache)


def test_server_hbm_budget_serialises():
    """A budget that holds one run at a time still answers every query."""
    g = pb.rmat(10, 16 * 2 ** 10, 0.57, 0.19, 0.19, seed=4)
    n = g.n
    one_run = 8 * n + 4 * ((n + 63) // 64) * 8 + 2 * 8 * n * 8   # single-source workspace
    cache = {}
    with pb.Server(0, slots=4, max_batch=1, hbm_budget=one_run + 1) as srv:
        gid = srv.add_graph(g)
        queries = {srv.submit(gid, s): (gid, s) for s in range(0, 1000, 37)}
        got = srv.drain(queries)
    check_results([g], got, queries, cache)


def test_server_rejects_bad_queries():
    g = pb.rmat(8, 1000, 0.57, 0.19, 0.19, seed=2)
    with pb.Server(0) as srv:
        gid = srv.add_graph(g)
        with pytest.raises(pb.PbbsError):
            srv.submit(gid, g.n)
        with pytest.raises(pb.PbbsError):
            srv.submit(gid + 1, 0)
    with pb.Server(0, hbm_budget=1024) as srv:
        gid = srv.add_graph(g)
        with pytest.raises(pb.PbbsError):
            srv.submit(gid, 0)
    with pytest.raises(pb.PbbsError):
        pb.Server(0, max_batch=65)
