"""The serving workload at its full size: 8 resident symmetric RMAT graphs of
2^16 to 2^22 vertices (16 draws per vertex, a=.57 b=.19 c=.19; the two
2^22 graphs from different seeds), a burst and then a seeded Poisson stream
of single-source queries on random graphs and random non-isolated sources.

Every query on a graph of up to 2^20 vertices is compared with the oracle
(deterministicBFS restated, parents; level_number, depths) bit for bit.
Queries on the 2^21 and 2^22 graphs are validated by the GPU checker (the
BFSCheck rule, the max-parent rule and depth == level), and a seeded sample
of them is compared with the oracle too."""
import threading
import time

import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb

pytestmark = pytest.mark.gpu

SERVE_GRAPHS = [(16, 1), (17, 1), (18, 1), (19, 1), (20, 1), (21, 1), (22, 1), (22, 2)]
ORACLE_ALL_UP_TO = 20       # scales whose every query is compared with the oracle
ORACLE_SAMPLE = 4           # oracle comparisons per larger graph


def say(*a):
    print("[case]", *a, flush=True)


def oracle_pd(g, s, cache):
    key = (id(g), s)
    if key not in cache:
        p, _ = oracle.det_bfs(g.offsets, g.adj, s, threads=16)
        cache[key] = (p, oracle.level_number(p))
    return cache[key]


def test_server_at_size_poisson_stream():
    t0 = time.time()
    graphs = [pb.rmat(sc, 16 * 2 ** sc, 0.57, 0.19, 0.19, seed=sd) for sc, sd in SERVE_GRAPHS]
    nz = [np.flatnonzero(g.degrees() > 0) for g in graphs]
    say(f"generated {len(graphs)} graphs in {time.time() - t0:.1f} s:",
        [(g.n, g.m) for g in graphs])
    rng = np.random.default_rng(2024)
    Q_BURST, Q_STREAM, QPS = 96, 320, 500.0
    Q = Q_BURST + Q_STREAM
    gsel = rng.integers(0, len(graphs), Q)
    srcs = [int(nz[gi][rng.integers(0, nz[gi].size)]) for gi in gsel]
    gaps = rng.exponential(1.0 / QPS, Q_STREAM)
    big = [gi for gi, (sc, _) in enumerate(SERVE_GRAPHS) if sc > ORACLE_ALL_UP_TO]
    sample = set()
    for gi in big:
        qs = [q for q in range(Q) if gsel[q] == gi]
        sample |= set(rng.choice(qs, min(ORACLE_SAMPLE, len(qs)), replace=False).tolist())

    # the GPU checker runs on copies of the large graphs in a context of its own
    ctx = pb.Context(0)
    check_dg = {gi: ctx.upload(graphs[gi]) for gi in big}
    cache = {}
    checked = {"oracle": 0, "gpu_checker": 0}
    lat, batches = [], []

    def verify(q, r, p, d):
        gi = int(gsel[q])
        g = graphs[gi]
        assert r["status"] == 0, r
        assert r["graph"] == ids[gi] and r["source"] == srcs[q], (r, q)
        if SERVE_GRAPHS[gi][0] <= ORACLE_ALL_UP_TO or q in sample:
            wp, wd = oracle_pd(g, srcs[q], cache)
            if not np.array_equal(p, wp):
                bad = np.flatnonzero(p != wp)[:8]
                raise AssertionError(f"query {q} graph {gi} source {srcs[q]} batch {r['batch']}: "
                                     f"parent differs at {bad}: {p[bad]} vs {wp[bad]}")
            assert np.array_equal(d, wd), f"query {q}: depth differs"
            checked["oracle"] += 1
        if gi in check_dg:
            c = check_dg[gi].check(srcs[q], p, d)
            assert c["valid"] == 1 and c["max_rule_bad"] == 0 and c["depth_bad"] == 0, (q, c)
            checked["gpu_checker"] += 1
        lat.append(r["latency_ms"])
        batches.append(r["batch"])

    with pb.Server(0) as srv:
        ids = [srv.add_graph(g) for g in graphs]
        say("graphs resident; burst of", Q_BURST)
        tickets = {}
        lock = threading.Lock()
        for q in range(Q_BURST):
            tickets[srv.submit(ids[gsel[q]], srcs[q])] = q
        got = srv.drain(list(tickets))
        for t, (r, p, d) in got.items():
            verify(tickets[t], r, p, d)
        say(f"burst verified; max batch {max(batches)}")
        assert max(batches) > 1, "a burst over 8 graphs should merge queries of one graph"

        def client():
            t = time.perf_counter()
            for i in range(Q_STREAM):
                q = Q_BURST + i
                t += gaps[i]
                dt = t - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                with lock:
                    tickets[srv.submit(ids[gsel[q]], srcs[q])] = q

        th = threading.Thread(target=client)
        th.start()
        done = 0
        deadline = time.monotonic() + 300
        while done < Q_STREAM:
            assert time.monotonic() < deadline, f"{Q_STREAM - done} stream queries unfinished"
            for r, p, d in srv.poll(wait_ms=50):
                with lock:
                    q = tickets[r["ticket"]]
                verify(q, r, p, d)
                done += 1
        th.join()
        ld = srv.load()
        assert ld["queued"] == 0 and ld["running"] == 0, ld
    for dg in check_dg.values():
        dg.free()
    ctx.close()
    lat = np.array(lat)
    say(f"{Q} queries: oracle-checked {checked['oracle']}, GPU-checked {checked['gpu_checker']}; "
        f"p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms; "
        f"mean batch {np.mean(batches):.2f}")
    assert checked["oracle"] >= sum(1 for q in range(Q)
                                    if SERVE_GRAPHS[gsel[q]][0] <= ORACLE_ALL_UP_TO)
    assert checked["gpu_checker"] == sum(1 for q in range(Q) if gsel[q] in check_dg)
