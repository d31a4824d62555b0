"""Pins the oracle before anything is compared against it.

The reference has no golden vectors for BFS; the pins are (a) the committed
known-answer fixtures (tests/golden, derived by formula per graph family),
(b) an independent brute force of the max-parent closed form, (c) the torus
closed form, (d) planted mistakes that the checker and the max rule must
reject.
"""
import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from conftest import GOLDEN, golden_cases


def load_golden(name):
    g = pb.read_graph_from_file(f"{GOLDEN}/{name}.adj")
    return (g, pb.read_int_seq_from_file(f"{GOLDEN}/{name}.parent"),
            pb.read_int_seq_from_file(f"{GOLDEN}/{name}.depth"))


@pytest.mark.parametrize("name,meta", golden_cases(), ids=[c[0] for c in golden_cases()])
def test_oracle_matches_golden(name, meta):
    g, wp, wd = load_golden(name)
    s = meta["source"]
    p, rounds = oracle.det_bfs(g.offsets, g.adj, s)
    assert np.array_equal(p, wp)
    ps, ds = oracle.serial_bfs(g.offsets, g.adj, s)
    assert np.array_equal(ps, wp) and np.array_equal(ds, wd)
    assert np.array_equal(oracle.level_number(p), wd)
    # rounds of deterministicBFS = levels reached by an expandable frontier
    assert rounds >= 1
    if meta["symmetric"]:
        assert oracle.check_bfs(g.offsets, g.adj, s, p) == -1
        assert oracle.max_rule_violations(g.offsets, g.adj, s, p, wd) == 0


def random_graph(seed, n, m, symmetric=True):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    # a mix of local and long edges, plus a few duplicates and self-loops
    v = np.where(rng.random(m) < 0.5, (u + rng.integers(-3, 4, m)) % n, rng.integers(0, n, m))
    return pb.from_edges(n, u, v, symmetrise=symmetric)


@pytest.mark.parametrize("seed", range(12))
def test_det_bfs_equals_brute_force(seed):
    n = [50, 300, 1000, 4000][seed % 4]
    g = random_graph(seed, n, 3 * n)
    rng = np.random.default_rng(100 + seed)
    for s in rng.integers(0, n, 3):
        p, _ = oracle.det_bfs(g.offsets, g.adj, int(s), threads=4)
        ps, ds = oracle.serial_bfs(g.offsets, g.adj, int(s))
        assert np.array_equal(p, ps)
        assert np.array_equal(oracle.level_number(p), ds)
        assert oracle.check_bfs(g.offsets, g.adj, int(s), p) == -1


@pytest.mark.parametrize("seed", range(4))
def test_det_bfs_directed_equals_brute_force(seed):
    g = random_graph(seed, 800, 2400, symmetric=False)
    p, _ = oracle.det_bfs(g.offsets, g.adj, 7)
    ps, ds = oracle.serial_bfs(g.offsets, g.adj, 7)
    assert np.array_equal(p, ps)
    assert np.array_equal(oracle.level_number(p), ds)


def test_det_bfs_thread_count_invariance():
    g = pb.rmat(13, 16 * 2 ** 13, 0.57, 0.19, 0.19, seed=3)
    ref, _ = oracle.det_bfs(g.offsets, g.adj, 1, threads=1)
    for t in (2, 3, 8):
        p, _ = oracle.det_bfs(g.offsets, g.adj, 1, threads=t)
        assert np.array_equal(p, ref)


@pytest.mark.parametrize("d", [3, 4, 7, 12])
def test_torus_closed_form_equals_brute_force(d):
    g = pb.grid3d(d)
    cp, cd = oracle.torus_closed_form(d)
    ps, ds = oracle.serial_bfs(g.offsets, g.adj, 0)
    assert np.array_equal(cp, ps) and np.array_equal(cd, ds)


def test_rmat_config1_oracle_consistency():
    # RMAT 2^16, 2^19 draws, (.5,.1,.1): the CPU reference configuration.
    g = pb.rmat(16, 2 ** 19, 0.5, 0.1, 0.1, seed=1)
    p, _ = oracle.det_bfs(g.offsets, g.adj, 0)
    ps, ds = oracle.serial_bfs(g.offsets, g.adj, 0)
    assert np.array_equal(p, ps)
    assert oracle.check_bfs(g.offsets, g.adj, 0, p) == -1
    assert oracle.max_rule_violations(g.offsets, g.adj, 0, p, ds) == 0


# ------------------------------------------------------- planted mistakes
def test_checker_rejects_planted_mistakes():
    g, wp, wd = load_golden("bintree31")
    assert oracle.check_bfs(g.offsets, g.adj, 0, wp) == -1
    bad = wp.copy(); bad[0] = 1                   # root does not point to itself
    assert oracle.check_bfs(g.offsets, g.adj, 0, bad) != -1
    bad = wp.copy(); bad[30] = 0                  # leaf claims level 1
    # its tree neighbour 14 (level 3) now sees a level-1 neighbour and fails first
    assert oracle.check_bfs(g.offsets, g.adj, 0, bad) == 14
    bad = wp.copy(); bad[29] = -1                 # reachable vertex left unreached
    assert oracle.check_bfs(g.offsets, g.adj, 0, bad) != -1
    bad = wp.copy(); bad[3] = 7; bad[7] = 3       # cycle
    assert oracle.check_bfs(g.offsets, g.adj, 0, bad) != -1
    assert oracle.check_bfs(g.offsets, g.adj, 0, wp[:-1]) == -2  # size mismatch
    g2, wp2, _ = load_golden("pairs12")
    bad = wp2.copy(); bad[8] = 6                  # unreachable vertex marked reached
    assert oracle.check_bfs(g2.offsets, g2.adj, 6, bad) != -1


def test_max_rule_rejects_min_parent():
    # The checker accepts any BFS tree; the max rule pins the choice.
    g, wp, wd = load_golden("layered11")
    lo = wp.copy()
    lo[6:10] = 1                                  # smallest level-1 neighbour instead
    assert oracle.check_bfs(g.offsets, g.adj, 0, lo) == -1
    assert oracle.max_rule_violations(g.offsets, g.adj, 0, lo, wd) == 4
    assert oracle.max_rule_violations(g.offsets, g.adj, 0, wp, wd) == 0


def test_level_number_rejects_cycle():
    with pytest.raises(ValueError):
        oracle.level_number(np.array([0, 2, 1], dtype=np.int32))


# ------------------------------------------------------------------- MIS
def _lex_first_mis_brute(n, edges):
    """Every subset of vertices; keep the maximal independent ones and return
    the lexicographically first (vertex 0 in before out, then vertex 1, ...)."""
    adj = [set() for _ in range(n)]
    for u, v in edges:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    best = None
    for mask in range(1 << n):
        S = [(mask >> i) & 1 for i in range(n)]
        if any(S[u] and S[v] for u in range(n) for v in adj[u]):
            continue
        if any(not S[v] and not any(S[u] for u in adj[v]) for v in range(n)):
            continue
        key = tuple(1 - x for x in S)   # smaller key: earlier vertices in the set
        if best is None or key < best[0]:
            best = (key, S)
    return np.array([1 if x else 2 for x in best[1]], dtype=np.int32)


@pytest.mark.parametrize("seed", range(12))
def test_mis_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 11))
    k = int(rng.integers(0, 2 * n + 1))
    u = rng.integers(0, n, k)
    v = rng.integers(0, n, k)
    g = pb.from_edges(n, u, v)
    want = _lex_first_mis_brute(n, list(zip(u.tolist(), v.tolist())))
    got = oracle.mis(g.offsets, g.adj)
    assert np.array_equal(got, want), (n, list(zip(u, v)))
    assert oracle.check_mis(g.offsets, g.adj, got) == -1


def test_mis_closed_forms():
    n = 101
    path = pb.from_edges(n, np.arange(n - 1), np.arange(1, n))
    assert np.array_equal(oracle.mis(path.offsets, path.adj),
                          np.where(np.arange(n) % 2 == 0, 1, 2))
    star0 = pb.from_edges(n, np.zeros(n - 1, np.int64), np.arange(1, n))
    assert np.array_equal(oracle.mis(star0.offsets, star0.adj), np.r_[1, np.full(n - 1, 2)])
    star_last = pb.from_edges(n, np.full(n - 1, n - 1), np.arange(n - 1))
    assert np.array_equal(oracle.mis(star_last.offsets, star_last.adj),
                          np.r_[np.full(n - 1, 1), 2])
    kn = 9
    iu, iv = np.triu_indices(kn, 1)
    comp = pb.from_edges(kn, iu, iv)
    assert np.array_equal(oracle.mis(comp.offsets, comp.adj), np.r_[1, np.full(kn - 1, 2)])
    for d in (4, 6, 10):
        t = pb.grid3d(d)
        assert np.array_equal(oracle.mis(t.offsets, t.adj), oracle.torus_mis_closed_form(d)), d


def test_check_mis_rejects_planted_mistakes():
    g = pb.rmat(9, 4000, 0.57, 0.19, 0.19, seed=3)
    f = oracle.mis(g.offsets, g.adj)
    assert oracle.check_mis(g.offsets, g.adj, f) == -1
    # a member's neighbour also in the set
    v = int(np.flatnonzero((f == 1) & (g.degrees() > 0))[0])
    u = int(g.adj[g.offsets[v]])
    bad = f.copy()
    bad[u] = 1
    assert oracle.check_mis(g.offsets, g.adj, bad) >= 0
    # a member dropped: its neighbours (or itself) lose their witness
    bad = f.copy()
    bad[v] = 2
    assert oracle.check_mis(g.offsets, g.adj, bad) >= 0


# ---- spanning forest (incrementalST under speculative_for's schedule) ------

def _st_edges(seed, n, m, loops=True):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    if not loops:
        v = np.where(u == v, (v + 1) % n, v).astype(np.int32)
    return u, v


def test_st_schedule_decides_the_forest():
    """Hand-derived: with m = 1200 the first round holds edges 0-2 (size
    (1200 / 100 + 1) / 4 = 3).  Edges 0 = (1,2) and 1 = (0,2) both reserve
    R[2]; edge 0 wins and links 2 under 1.  Edge 2 = (0,1) reserves R[1] and
    links 1 under 0 in the same round, so edge 1 later finds one tree and is
    dropped.  Output R in vertex order: R[1] = 2, R[2] = 0.  A serial pass
    keeps edges 0 and 1 instead."""
    u = np.r_[[1, 0, 0], np.full(1197, 3)].astype(np.int32)
    v = np.r_[[2, 2, 1], np.full(1197, 3)].astype(np.int32)
    f, rounds = oracle.st(4, u, v)
    assert f.tolist() == [2, 0]
    assert oracle.serial_st(4, u, v).tolist() == [0, 1]
    # a round holds at most 1200 / 100 + 1 = 13 iterates
    assert oracle.check_st(4, u, v, f) == 0 and rounds >= -(-1200 // 13)


def test_st_closed_forms():
    # star with the hub last: edge 0 takes R[hub]; every other edge i then
    # joins leaf i to the tree of 0 through R[i]
    n = 2001
    f, _ = oracle.st(n, np.arange(n - 1), np.full(n - 1, n - 1))
    assert f.tolist() == list(range(1, n - 1)) + [0]
    # a path listed in order: edge i = (i, i+1) always holds R[i+1] alone
    f, _ = oracle.st(n, np.arange(n - 1), np.arange(1, n))
    assert f.tolist() == list(range(n - 1))
    # no edges at all; and m < 300, where the reference's schedule gives up
    f, _ = oracle.st(5, np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert f.tolist() == []
    f, rounds = oracle.st(5, np.zeros(299, np.int32), np.ones(299, np.int32))
    assert f is None and rounds > 100 + 200 * 100
    f, _ = oracle.st(5, np.zeros(300, np.int32), np.ones(300, np.int32))
    assert f.tolist() == [0]


@pytest.mark.parametrize("seed", range(6))
def test_st_is_a_spanning_forest(seed):
    """The schedule's forest spans what the serial forest spans (same size,
    n - components) and has no cycle; it is not always the serial one."""
    n, m = [(50, 400), (300, 2000), (2000, 5000), (5000, 40000), (100, 3000), (4000, 4000)][seed]
    u, v = _st_edges(seed, n, m, loops=seed % 2 == 0)
    f, _ = oracle.st(n, u, v)
    s = oracle.serial_st(n, u, v)
    assert f.size == s.size and np.unique(f).size == f.size
    assert oracle.check_st(n, u, v, f) == 0 and oracle.check_st(n, u, v, s) == 0
    # components by scipy agree with n - |forest|
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    k, _ = connected_components(coo_matrix((np.ones(m), (u, v)), shape=(n, n)), directed=False)
    assert f.size == n - k


def test_check_st_rejects_planted_mistakes():
    u, v = _st_edges(7, 500, 3000)
    f, _ = oracle.st(500, u, v)
    assert oracle.check_st(500, u, v, f) == 0
    assert oracle.check_st(500, u, v, f[:-1]) == 1            # one edge short
    # swap a forest edge for a non-forest edge closing a cycle
    inf = set(f.tolist())
    extra = next(i for i in range(3000) if i not in inf and u[i] != v[i])
    bad = np.r_[f[:-1], extra].astype(np.int32)
    s = oracle.serial_st(500, u[bad], v[bad])
    assert oracle.check_st(500, u, v, bad) in (0, 2) and (s.size < bad.size) == (
        oracle.check_st(500, u, v, bad) == 2)
    dup = np.r_[f[:-1], f[0]].astype(np.int32)                 # a repeated edge is a cycle
    assert oracle.check_st(500, u, v, dup) == 2
    out = f.copy()
    out[0] = 3000                                              # not an edge id
    assert oracle.check_st(500, u, v, out) == 3
