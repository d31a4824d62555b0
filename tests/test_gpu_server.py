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
        ld = srv.load()
        assert ld["queued"] == 0 and ld["running"] == 0 and ld["hbm_in_use"] > 0, ld
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


def hip_free_bytes():
    """Free device memory as the HIP runtime reports it (hipMemGetInfo, found
    through the library's own HIP runtime dependency)."""
    import ctypes as C
    from pbbsbench_amd import _lib
    fn = _lib.lib().hipMemGetInfo
    fn.restype = C.c_int
    free, total = C.c_size_t(), C.c_size_t()
    assert fn(C.byref(free), C.byref(total)) == 0
    return free.value


def test_server_hbm_budget_bounds_device_memory(ctx):
    """A budget that holds the graphs plus one batched run of 8 sources: the
    server's charge never exceeds it, queries are still all answered, and
    the device memory the server takes stays within the budget (plus each
    context's fixed control blocks)."""
    graphs = [pb.rmat(16, 16 * 2 ** 16, 0.57, 0.19, 0.19, seed=4),
              pb.rmat(15, 16 * 2 ** 15, 0.57, 0.19, 0.19, seed=5)]
    # what the server will charge for each graph: the graph and its search data
    gbytes = 0
    for g in graphs:
        dg = ctx.upload(g)
        dg.prepare()
        gbytes += dg.info["device_bytes"]
        dg.free()
    n = graphs[0].n
    batched8 = 5 * n * 8 + 4 * ((n + 63) // 64) * 8 + 64 * n + 2 * 8 * n * 8 + 2 * 8 * n * 4
    budget = gbytes + batched8 + (1 << 20)
    slots = 4
    cache = {}
    free0 = hip_free_bytes()
    peak_charge, peak_drop = 0, 0
    with pb.Server(0, slots=slots, max_batch=8, hbm_budget=budget) as srv:
        ids = [srv.add_graph(g) for g in graphs]
        rng = np.random.default_rng(3)
        queries = {}
        for i in range(160):
            gi = int(rng.integers(0, 2))
            s = int(rng.integers(0, graphs[gi].n))
            queries[srv.submit(ids[gi], s, depth=(i % 2 == 0))] = (gi, s)
            if i % 16 == 0:
                peak_charge = max(peak_charge, srv.load()["hbm_in_use"])
                peak_drop = max(peak_drop, free0 - hip_free_bytes())
        got = {}
        while len(got) < len(queries):
            for r, p, d in srv.poll(wait_ms=20):
                got[r["ticket"]] = (r, p, d)
            peak_charge = max(peak_charge, srv.load()["hbm_in_use"])
            peak_drop = max(peak_drop, free0 - hip_free_bytes())
    say(f"budget {budget}, peak charge {peak_charge}, peak device drop {peak_drop}")
    assert peak_charge <= budget
    # contexts' fixed blocks (control state, per-workgroup counters, argument
    # blocks, captured launch graphs): at most 16 MiB each, server main + slots
    assert peak_drop <= budget + (slots + 1) * (16 << 20), (peak_drop, budget)
    check_results(graphs, got, {t: (ids[gi], s) for t, (gi, s) in queries.items()}, cache)


def test_server_rejects_bad_queries():
    g = pb.rmat(8, 1000, 0.57, 0.19, 0.19, seed=2)
    with pb.Server(0) as srv:
        gid = srv.add_graph(g)
        with pytest.raises(pb.PbbsError):
            srv.submit(gid, g.n)
        with pytest.raises(pb.PbbsError):
            srv.submit(gid + 1, 0)
    with pb.Server(0, hbm_budget=1024) as srv:   # the graph itself does not fit
        with pytest.raises(pb.PbbsError):
            srv.add_graph(g)
    with pytest.raises(pb.PbbsError):
        pb.Server(0, max_batch=65)
