"""Spanning forest on the GPU (pbbs_spanning_forest) against the oracle.

The answer is incrementalST's (spanningForest/incrementalST/ST.C:35-71): the
forest its speculative_for schedule (common/speculative_for.h:30-100,
granularity 100) builds, listed by R in vertex order.  The GPU runs the same
schedule, so the edge ids must equal the oracle's restatement bit for bit,
round count included; every result also passes STCheck's rule
(bench/STCheck.C:36-61).  Cases: the hand-derived round where the schedule
and a serial pass disagree, a hub-last star and an in-order path (closed
forms), random edge lists with self-loops and repeated edges, edges from
RMAT, torus and random-local graphs in shuffled order, edges and forest in
HBM, the reference's refusals (m < 300, ids outside [0, n)), and the
rMatGraph_E_12 shape at 2^22."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from pbbsbench_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def graph_edges(g, seed):
    """One (u, v) per undirected edge of a symmetric CSR (u < v), shuffled."""
    off = np.asarray(g.offsets)
    src = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(off))
    keep = src < g.adj
    u, v = src[keep], g.adj[keep]
    p = np.random.default_rng(seed).permutation(u.size)
    return u[p].astype(np.int32), v[p].astype(np.int32)


def run_st(ctx, n, u, v):
    say(f"spanning forest n={n} m={len(u)}")
    f, st = ctx.spanning_forest(n, u, v)
    f2, st2 = ctx.spanning_forest(n, u, v)
    want, rounds = oracle.st(n, u, v)
    assert np.array_equal(f, want)
    assert np.array_equal(f2, f)
    assert st["rounds"] == rounds == st2["rounds"]
    assert st["edges"] == f.size
    assert oracle.check_st(n, u, v, f) == 0
    return f, st


def test_st_round_where_schedule_and_serial_differ(ctx):
    u = np.r_[[1, 0, 0], np.full(1197, 3)].astype(np.int32)
    v = np.r_[[2, 2, 1], np.full(1197, 3)].astype(np.int32)
    f, _ = run_st(ctx, 4, u, v)
    assert f.tolist() == [2, 0]


def test_st_closed_forms(ctx):
    n = 2001
    f, _ = run_st(ctx, n, np.arange(n - 1), np.full(n - 1, n - 1))
    assert f.tolist() == list(range(1, n - 1)) + [0]
    f, _ = run_st(ctx, n, np.arange(n - 1), np.arange(1, n))
    assert f.tolist() == list(range(n - 1))
    f, st = ctx.spanning_forest(7, np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert f.size == 0 and st["rounds"] == 0


@pytest.mark.parametrize("seed,n,m", [(1, 300, 1000), (2, 5000, 20000), (3, 100000, 60000),
                                      (4, 1000, 200000), (5, 2, 5000)])
def test_st_random_edges(ctx, seed, n, m):
    """Self-loops, repeated edges, many components or one dense one."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    run_st(ctx, n, u, v)


def test_st_graph_shapes(ctx):
    for name, g, seed in (("rmat16", pb.rmat(16, 12 * 2 ** 16, 0.55, 0.125, 0.125, seed=3), 1),
                          ("torus32", pb.grid3d(32), 2),
                          ("randlocal", pb.randlocal(100_000, 10, seed=4), 3)):
        u, v = graph_edges(g, seed)
        say(name)
        run_st(ctx, g.n, u, v)
        # the same edges in id order: a different schedule, still the oracle's
        o = np.lexsort((v, u))
        run_st(ctx, g.n, u[o], v[o])


def test_st_refusals(ctx):
    lib = _lib.lib()
    u = np.zeros(299, np.int32)
    v = np.ones(299, np.int32)
    with pytest.raises(pb.PbbsError, match="gives up"):
        ctx.spanning_forest(2, u, v)
    u = np.zeros(1000, np.int32)
    v = np.ones(1000, np.int32)
    v[500] = 9
    with pytest.raises(pb.PbbsError, match="outside"):
        ctx.spanning_forest(5, u, v)
    cnt = C.c_int64()
    rc = lib.pbbs_spanning_forest(ctx._h, 5, 1000, u.ctypes.data, u.ctypes.data, 0x80, None,
                                  C.byref(cnt), None)
    assert rc == -1 and b"unknown flags" in lib.pbbs_last_error()
    # the context still works after the refusals
    run_st(ctx, 5, np.r_[np.arange(4), np.zeros(996)].astype(np.int32),
           np.r_[np.arange(1, 5), np.zeros(996)].astype(np.int32))


def test_st_device_edges_and_output(ctx):
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    g = pb.rmat(14, 12 * 2 ** 14, 0.55, 0.125, 0.125, seed=6, permute_seed=2)
    u, v = graph_edges(g, 5)
    bufs = [C.c_void_p() for _ in range(3)]
    for b, nbytes in zip(bufs, (u.nbytes, v.nbytes, g.n * 4)):
        assert hip.hipMalloc(C.byref(b), nbytes) == 0
    try:
        assert hip.hipMemcpy(bufs[0], u.ctypes.data, u.nbytes, 1) == 0   # host -> device
        assert hip.hipMemcpy(bufs[1], v.ctypes.data, v.nbytes, 1) == 0
        say("spanning forest, device edges and output")
        k, st = ctx.spanning_forest_device(g.n, u.size, bufs[0].value, bufs[1].value,
                                           bufs[2].value)
        got = np.empty(k, np.int32)
        assert hip.hipMemcpy(got.ctypes.data, bufs[2], k * 4, 2) == 0    # device -> host
        want, rounds = oracle.st(g.n, u, v)
        assert np.array_equal(got, want) and st["rounds"] == rounds
        # a smaller edge array next, through the same buffers
        k2, _ = ctx.spanning_forest_device(g.n, 5000, bufs[0].value, bufs[1].value,
                                           bufs[2].value)
        got2 = np.empty(k2, np.int32)
        assert hip.hipMemcpy(got2.ctypes.data, bufs[2], k2 * 4, 2) == 0
        assert np.array_equal(got2, oracle.st(g.n, u[:5000], v[:5000])[0])
    finally:
        for b in bufs:
            hip.hipFree(b)


def test_st_rmat22_edges():
    """rMatGraph_E_12's shape at 2^22: 12 * 2^22 draws (a=.55, b=c=.125),
    one entry per undirected edge, shuffled."""
    with pb.Context(0) as c:
        g = pb.rmat(22, 12 * 2 ** 22, 0.55, 0.125, 0.125, seed=1, permute_seed=2)
        u, v = graph_edges(g, 7)
        del g
        f, st = run_st(c, 2 ** 22, u, v)
        say(f"rounds={st['rounds']} edges={st['edges']} ms={st['device_ms']:.2f}")


def test_graph_edges_on_gpu(ctx):
    """pbbs_graph_edges: row order equals the CSR's upper entries; a seeded
    shuffle is a permutation of them, the same for the same seed; the forest
    of the device edges equals the oracle's on the downloaded ones."""
    g = pb.rmat(12, 12 * 2 ** 12, 0.55, 0.125, 0.125, seed=4, permute_seed=5)
    dg = ctx.upload(g)
    off = np.asarray(g.offsets)
    src = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(off))
    keep = src < g.adj
    want_u, want_v = src[keep], g.adj[keep]
    es = []
    try:
        say("graph edges: row order and two shuffles")
        for seed in (0, 9, 9, 10):
            es.append(dg.edges_device(seed))
        u0, v0 = es[0].download()
        assert np.array_equal(u0, want_u) and np.array_equal(v0, want_v)
        u1, v1 = es[1].download()
        u2, v2 = es[2].download()
        u3, v3 = es[3].download()
        assert np.array_equal(u1, u2) and np.array_equal(v1, v2)
        assert not np.array_equal(u1, u0) and not np.array_equal(u1, u3)
        k = lambda a, b: np.sort(a.astype(np.int64) << 32 | b.astype(np.int64))
        assert np.array_equal(k(u1, v1), k(u0, v0)) and np.array_equal(k(u3, v3), k(u0, v0))
        cnt, st = ctx.spanning_forest_device(g.n, es[1].m, es[1].u, es[1].v)
        f, _ = ctx.spanning_forest(g.n, u1, v1)
        want, rounds = oracle.st(g.n, u1, v1)
        assert cnt == f.size and np.array_equal(f, want) and st["rounds"] == rounds
    finally:
        for e in es:
            e.free()
        dg.free()


def test_st_any_forest(ctx):
    """PBBS_ST_ANY_FOREST (ndST's contract, spanningForest/ndST/ST.C): some
    spanning forest, checked by STCheck's rule (bench/STCheck.C:36-61) and by
    its size against the serial forest; ids are listed by hooked root, so
    they are distinct and each hooks a different vertex."""
    rng = np.random.default_rng(12)
    cases = [(50, rng.integers(0, 50, 120), rng.integers(0, 50, 120)),       # m < 300 is fine here
             (5000, rng.integers(0, 5000, 40000), rng.integers(0, 5000, 40000)),
             (2, np.zeros(5000, np.int64), np.ones(5000, np.int64))]
    g = pb.rmat(16, 12 * 2 ** 16, 0.55, 0.125, 0.125, seed=3, permute_seed=1)
    u, v = graph_edges(g, 4)
    cases.append((g.n, u, v))
    t = pb.grid3d(24)
    u, v = graph_edges(t, 5)
    cases.append((t.n, u, v))
    for n, u, v in cases:
        u = np.asarray(u, np.int32)
        v = np.asarray(v, np.int32)
        say(f"any forest n={n} m={u.size}")
        f, st = ctx.spanning_forest(n, u, v, any_forest=True)
        assert oracle.check_st(n, u, v, f) == 0
        assert f.size == oracle.serial_st(n, u, v).size == st["edges"]
        assert np.unique(f).size == f.size
    with pytest.raises(pb.PbbsError, match="outside"):
        ctx.spanning_forest(5, np.zeros(10, np.int32), np.full(10, 7, np.int32), any_forest=True)


def test_st_any_forest_rmat22():
    with pb.Context(0) as c:
        g = pb.rmat(22, 12 * 2 ** 22, 0.55, 0.125, 0.125, seed=1, permute_seed=2)
        u, v = graph_edges(g, 7)
        del g
        say("any forest rmat22")
        f, st = c.spanning_forest(2 ** 22, u, v, any_forest=True)
        assert oracle.check_st(2 ** 22, u, v, f) == 0
        say(f"edges={st['edges']} ms={st['device_ms']:.2f}")


def test_st_full_size_rmat24_edges():
    """The spanning-forest bench input at full size (190.6M shuffled RMAT
    edges, built in HBM): ids and round count equal the oracle's."""
    with pb.Context(0) as c:
        dg = c.gen_rmat(24, 12 * 16_000_000, 0.55, 0.125, 0.125, seed=1, permute_seed=2)
        n = dg.n
        e = dg.edges_device(shuffle_seed=3)
        dg.free()
        try:
            say(f"spanning forest full size m={e.m}")
            cnt, st = c.spanning_forest_device(n, e.m, e.u, e.v)
            u, v = e.download()
        finally:
            e.free()
        f, st2 = c.spanning_forest(n, u, v)
        want, rounds = oracle.st(n, u, v)
        assert cnt == f.size and st["rounds"] == st2["rounds"] == rounds
        assert np.array_equal(f, want)
