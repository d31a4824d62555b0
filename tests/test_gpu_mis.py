"""Maximal independent set on the GPU (pbbs_mis) against the oracle.

The set is incrementalMIS's (maximalIndependentSet/incrementalMIS/MIS.C:35-66):
v is in it unless a lower-numbered neighbour is, so it is one fixed set per
graph and the GPU's flags must equal the oracle's bit for bit; every result is
also run through MISCheck's rule (bench/MISCheck.C:37-63) and repeated runs
must agree.  Cases: the golden graphs, RMAT both offset widths, ordered and
relabelled tori (the even torus has a closed form), stars with the hub first
and last, a path (one decision per round: the longest chain), unsorted rows,
a directed graph (rows are out-neighbours, as in the reference), a graph with
no edges, and the benchmark's rMatGraph_JR_12 at full size."""
import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN, golden_cases

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def run_mis(ctx, g, wide=False, assume_symmetric=False):
    say(f"mis n={g.n} m={g.m} wide={wide}")
    dg = ctx.upload(g, wide_offsets=wide, assume_symmetric=assume_symmetric)
    symmetric = dg.info["symmetric"]
    try:
        flags, st = dg.mis()
        f2, st2 = dg.mis()
    finally:
        dg.free()
    want = oracle.mis(g.offsets, g.adj)
    assert np.array_equal(flags, want.astype(np.int32))
    assert np.array_equal(f2, flags)
    if symmetric:   # MISCheck's rule holds on symmetric graphs only
        assert oracle.check_mis(g.offsets, g.adj, flags) == -1
    assert st["members"] == int((flags == 1).sum()) == st2["members"]
    assert st["rounds"] >= 1 and st2["rounds"] >= 1
    return flags, st


@pytest.mark.parametrize("name,meta", golden_cases(), ids=[c[0] for c in golden_cases()])
def test_mis_golden(ctx, name, meta):
    g = pb.read_graph_from_file(f"{GOLDEN}/{name}.adj")
    run_mis(ctx, g, assume_symmetric=meta["symmetric"])


@pytest.mark.parametrize("wide", [False, True])
def test_mis_rmat(ctx, wide):
    g = pb.rmat(14, 12 * 2 ** 14, 0.55, 0.125, 0.125, seed=3)
    run_mis(ctx, g, wide=wide)
    run_mis(ctx, pb.rmat(14, 12 * 2 ** 14, 0.55, 0.125, 0.125, seed=3, permute_seed=4), wide=wide)


def test_mis_torus():
    with pb.Context(0) as c:
        g = pb.grid3d(10)
        flags, _ = run_mis(c, g)
        assert np.array_equal(flags, oracle.torus_mis_closed_form(10).astype(np.int32))
        run_mis(c, pb.grid3d(9))                 # odd side: the wrap edges break the pattern
        run_mis(c, pb.permute(pb.grid3d(24), 5))  # 3Dgrid_JR's shape


def test_mis_star_path_empty(ctx):
    n = 50_001
    hub_first = pb.from_edges(n, np.zeros(n - 1, np.int64), np.arange(1, n))
    f, _ = run_mis(ctx, hub_first)
    assert f[0] == 1 and (f[1:] == 2).all()
    hub_last = pb.from_edges(n, np.full(n - 1, n - 1, np.int64), np.arange(n - 1))
    f, _ = run_mis(ctx, hub_last)
    assert f[-1] == 2 and (f[:-1] == 1).all()
    k = 3000
    path = pb.from_edges(k, np.arange(k - 1), np.arange(1, k))
    f, st = run_mis(ctx, path)
    assert (f[0::2] == 1).all() and (f[1::2] == 2).all()
    assert st["rounds"] >= k // 2    # each round can only settle the next link of the chain
    empty = pb.HostGraph(np.zeros(1001, np.int64), np.zeros(0, np.int32))
    f, st = run_mis(ctx, empty)
    assert (f == 1).all() and st["rounds"] == 1


def test_mis_unsorted_rows(ctx):
    g = pb.permute(pb.rmat(12, 16 * 2 ** 12, 0.57, 0.19, 0.19, seed=9), 3)
    adj = g.adj.copy()
    for v in range(g.n):
        adj[g.offsets[v]:g.offsets[v + 1]] = adj[g.offsets[v]:g.offsets[v + 1]][::-1]
    h = pb.HostGraph(g.offsets, adj)
    dg = ctx.upload(h, assume_symmetric=True)
    assert not dg.info["sorted"]
    dg.free()
    run_mis(ctx, h, assume_symmetric=True)


def test_mis_directed(ctx):
    rng = np.random.default_rng(5)
    u = rng.integers(0, 3000, 20000)
    v = rng.integers(0, 3000, 20000)
    run_mis(ctx, pb.from_edges(3000, u, v, symmetrise=False))


def test_mis_device_output_and_smaller_graph(ctx):
    """Flags left in HBM equal the host copy; a larger graph then a smaller
    one reuse the workspace."""
    import ctypes as C
    import os
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    big = pb.permute(pb.randlocal(200_000, 10, seed=2), 9)
    small = pb.grid3d(6)
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), big.n * 4) == 0
    try:
        for g in (big, small):
            dg = ctx.upload(g)
            try:
                say(f"mis device output n={g.n}")
                st = dg.mis_device(buf.value)
                host, _ = dg.mis()
            finally:
                dg.free()
            dev = np.empty(g.n, dtype=np.int32)
            assert hip.hipMemcpy(dev.ctypes.data, buf, g.n * 4, 2) == 0   # device -> host
            assert np.array_equal(dev, host)
            assert np.array_equal(host, oracle.mis(g.offsets, g.adj).astype(np.int32))
            assert st["members"] == int((host == 1).sum())
    finally:
        hip.hipFree(buf)


def test_mis_rejects_bad_flags(ctx):
    import ctypes as C
    from pbbsbench_amd import _lib
    g = pb.grid3d(4)
    dg = ctx.upload(g)
    try:
        out = np.empty(g.n, np.int32)
        rc = _lib.lib().pbbs_mis(ctx._h, dg._h, out.ctypes.data_as(C.c_void_p), 0x80, None)
        assert rc == -1 and b"unknown flags" in _lib.lib().pbbs_last_error()
    finally:
        dg.free()


def test_mis_full_size():
    """rMatGraph_JR_12_16000000's shape: 2^24 vertices, 12 * 16M edge draws
    (a=.55, b=c=.125), relabelled; the flags equal the oracle's."""
    with pb.Context(0) as c:
        say("mis full size: rmat24 x12 relabelled")
        dg = c.gen_rmat(24, 12 * 16_000_000, 0.55, 0.125, 0.125, seed=1, permute_seed=2)
        try:
            flags, st = dg.mis()
            g = dg.download()
            r = dg.check_mis(flags)
        finally:
            dg.free()
        say(f"rounds={st['rounds']} members={st['members']} ms={st['device_ms']:.2f}")
        assert r["valid"] == 1, r
        want = oracle.mis(g.offsets, g.adj)
        assert np.array_equal(flags, want.astype(np.int32))


def test_check_mis_on_gpu(ctx):
    """pbbs_check_mis accepts the set and rejects planted faults at the vertex
    the oracle's MISCheck restatement names."""
    g = pb.permute(pb.rmat(13, 12 * 2 ** 13, 0.55, 0.125, 0.125, seed=8), 6)
    dg = ctx.upload(g)
    try:
        say("check_mis rmat13")
        flags, _ = dg.mis()
        r = dg.check_mis(flags)
        assert r["valid"] == 1 and r["first_bad"] == -1
        off = np.asarray(g.offsets)
        members = np.flatnonzero(flags == 1)
        # a member's neighbour joins too: bad edge at both ends
        v = int(members[np.argmax(np.diff(off)[members] > 0)])
        u = int(g.adj[off[v]])
        f2 = flags.copy()
        f2[u] = 1
        r = dg.check_mis(f2)
        want = oracle.check_mis(g.offsets, g.adj, f2)
        assert r["valid"] == 0 and r["first_bad"] == want <= min(u, v)
        assert r["bad_edges"] >= 2
        # a member leaves: it (and maybe neighbours) have no member neighbour
        f3 = flags.copy()
        f3[v] = 2
        r = dg.check_mis(f3)
        assert r["valid"] == 0 and r["first_bad"] == oracle.check_mis(g.offsets, g.adj, f3)
        assert r["bad_vertices"] >= 1 and r["bad_edges"] == 0
        # the rule on device input
        import ctypes as C
        import os
        hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                  "libamdhip64.so"))
        hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        hip.hipFree.argtypes = [C.c_void_p]
        buf = C.c_void_p()
        assert hip.hipMalloc(C.byref(buf), g.n * 4) == 0
        try:
            dg.mis_device(buf.value)
            r = dg.check_mis_device(buf.value)
            assert r["valid"] == 1
        finally:
            hip.hipFree(buf)
    finally:
        dg.free()


def test_mis_full_size_randlocal_and_grid():
    """randLocalGraph_JR_10_20000000's and 3Dgrid_JR_64000000's shapes at full
    size (relabelled, rows sorted): the flags equal the oracle's, and the GPU
    MISCheck accepts them."""
    with pb.Context(0) as c:
        for name, g in (("randlocal 20M", pb.permute(pb.randlocal(20_000_000, 10, seed=1), 3)),
                        ("3D torus 400^3", pb.permute(pb.grid3d(400), 5))):
            say(f"mis full size: {name} n={g.n} m={g.m}")
            dg = c.upload(g, assume_symmetric=True)
            try:
                flags, st = dg.mis()
                r = dg.check_mis(flags)
            finally:
                dg.free()
            say(f"rounds={st['rounds']} ms={st['device_ms']:.2f}")
            assert r["valid"] == 1, r
            assert np.array_equal(flags, oracle.mis(g.offsets, g.adj))
            del g
