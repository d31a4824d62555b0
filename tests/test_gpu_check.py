"""GPU checker (pbbs_check_bfs) against the oracle's checkBFS restatement and
max rule: same verdict and same first failing vertex on correct outputs and on
the planted mistakes the oracle's own tests use; same max-rule count."""
import os

import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def golden(name):
    g = pb.read_graph_from_file(os.path.join(GOLDEN, f"{name}.adj"))
    wp = pb.read_int_seq_from_file(os.path.join(GOLDEN, f"{name}.parent"))
    wd = pb.read_int_seq_from_file(os.path.join(GOLDEN, f"{name}.depth"))
    return g, wp, wd


def agree(dg, g, s, parent, depth=None):
    """GPU verdict vs oracle; returns the GPU result."""
    r = dg.check(s, parent, depth)
    want = oracle.check_bfs(g.offsets, g.adj, s, parent)
    if want == -2:                       # cycle (sizes always match here)
        assert r["cycle"] == 1 and r["valid"] == 0, r
        return r
    assert r["cycle"] == 0
    assert (r["valid"] == 1) == (want == -1), (r, want)
    assert r["first_bad"] == want, (r, want)
    if r["valid"] or r["bad"]:
        if all(-1 <= p < g.n for p in parent):
            lv = oracle.level_number(parent)
            assert r["max_rule_bad"] == oracle.max_rule_violations(g.offsets, g.adj, s, parent,
                                                                   lv)
    return r


def test_check_accepts_correct_outputs(ctx):
    for name in ("bintree31", "layered11", "torus4", "comb10", "pairs12"):
        g, wp, wd = golden(name)
        s = int(np.flatnonzero(wp == np.arange(g.n))[0])
        say(name, s)
        dg = ctx.upload(g)
        r = agree(dg, g, s, wp, wd)
        assert r["valid"] == 1 and r["max_rule_bad"] == 0 and r["depth_bad"] == 0
        assert r["reached"] == int((wd < g.n).sum()) and r["max_level"] == int(wd[wd < g.n].max())
        dg.free()


def test_check_planted_mistakes(ctx):
    g, wp, wd = golden("bintree31")
    dg = ctx.upload(g)
    plants = []
    bad = wp.copy(); bad[0] = 1; plants.append(bad)                 # root not its own parent
    bad = wp.copy(); bad[30] = 0; plants.append(bad)                # leaf claims level 1
    bad = wp.copy(); bad[29] = -1; plants.append(bad)               # reachable left unreached
    bad = wp.copy(); bad[3] = 7; bad[7] = 3; plants.append(bad)     # cycle
    bad = wp.copy(); bad[5] = 31; plants.append(bad)                # parent outside [-1, n)
    bad = wp.copy(); bad[6] = -7; plants.append(bad)
    for i, p in enumerate(plants):
        say("plant", i)
        r = agree(dg, g, 0, p)
        assert r["valid"] == 0
    r = dg.check(0, wp, wd + (np.arange(g.n) == 9))                  # one depth off
    assert r["valid"] == 1 and r["depth_bad"] == 1
    dg.free()
    g2, wp2, _ = golden("pairs12")
    dg2 = ctx.upload(g2)
    bad = wp2.copy(); bad[8] = 6                                      # unreachable marked reached
    assert agree(dg2, g2, 6, bad)["valid"] == 0
    dg2.free()


def test_check_max_rule(ctx):
    g, wp, wd = golden("layered11")
    dg = ctx.upload(g)
    lo = wp.copy()
    lo[6:10] = 1                           # smallest level-1 neighbour instead of largest
    r = agree(dg, g, 0, lo, wd)
    assert r["valid"] == 1 and r["max_rule_bad"] == 4
    dg.free()


def test_check_directed_graph_skips_max_rule(ctx):
    """On a graph that is not symmetric a row lists out-neighbours, not the
    in-neighbours the max rule needs: the checker reports -1 for it, and
    the BFSCheck rule itself still matches the oracle's verdict."""
    g, wp, wd = golden("directed6")
    dg = ctx.upload(g)
    assert dg.info["symmetric"] == 0
    p, d, _ = dg.bfs(0)
    assert np.array_equal(p, wp) and np.array_equal(d, wd)
    r = dg.check(0, wp, wd)
    assert r["max_rule_bad"] == -1 and r["depth_bad"] == 0
    want = oracle.check_bfs(g.offsets, g.adj, 0, wp)
    assert (r["valid"] == 1) == (want == -1) and r["first_bad"] == want, (r, want)
    dg.free()


@pytest.mark.parametrize("wide", [False, True])
def test_check_large_and_batched(ctx, wide):
    """RMAT 2^14: a real BFS passes, a shuffled parent fails the max rule, and
    the rows of a batched output are checked in HBM."""
    g = pb.rmat(14, 16 * 2 ** 14, 0.57, 0.19, 0.19, seed=6)
    dg = ctx.upload(g, wide_offsets=wide)
    rng = np.random.default_rng(1)
    src = rng.choice(np.flatnonzero(g.degrees() > 0), 5, replace=False)
    p, d, _ = dg.bfs(int(src[0]))
    r = agree(dg, g, int(src[0]), p, d)
    assert r["valid"] and r["max_rule_bad"] == 0 and r["depth_bad"] == 0
    # a non-max parent that is still one level up: valid tree, max rule broken
    lv = oracle.level_number(p)
    v = next(v for v in range(g.n)
             if 1 < lv[v] < g.n and len({u for u in g.adj[g.offsets[v]:g.offsets[v + 1]]
                                         if lv[u] == lv[v] - 1}) > 1)
    alt = min(u for u in g.adj[g.offsets[v]:g.offsets[v + 1]] if lv[u] == lv[v] - 1)
    q = p.copy(); q[v] = alt
    r = agree(dg, g, int(src[0]), q)
    assert r["valid"] == 1 and r["max_rule_bad"] == 1
    # batched rows, checked where they are (the context's output buffers)
    dg.bfs_batch_device(src)
    pp, dp = ctx.outputs(batched=True)
    for i, s in enumerate(src):
        r = dg.check_device(int(s), pp + 4 * i * g.n, dp + 4 * i * g.n)
        assert r["valid"] and r["max_rule_bad"] == 0 and r["depth_bad"] == 0, (i, r)
    dg.bfs_device(int(src[1]))
    pp, dp = ctx.outputs()
    r = dg.check_device(int(src[1]), pp, dp)
    assert r["valid"] and r["max_rule_bad"] == 0 and r["depth_bad"] == 0
    dg.free()
