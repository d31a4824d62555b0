"""Host CSR construction and the seeded generators (product code, CPU only)."""
import numpy as np
import pytest

import pbbsbench_amd as pb


def rows_of(g):
    return np.repeat(np.arange(g.n, dtype=np.int64), g.degrees())


def assert_canonical_symmetric(g):
    """Sorted strictly ascending rows, no self-loops, symmetric."""
    r = rows_of(g)
    assert np.all(g.adj >= 0) and np.all(g.adj < g.n)
    assert not np.any(r == g.adj)
    key = r * g.n + g.adj
    assert np.all(np.diff(key) > 0)  # sorted by (u,v), no duplicates
    rkey = np.sort(g.adj.astype(np.int64) * g.n + r)
    assert np.array_equal(key, rkey)


def test_csr_from_edges_dedups_and_symmetrises():
    g = pb.from_edges(5, [0, 0, 1, 2, 2, 3], [1, 1, 0, 2, 3, 4])
    assert list(g.offsets) == [0, 1, 2, 3, 5, 6]
    assert list(g.adj) == [1, 0, 3, 2, 4, 3]
    d = pb.from_edges(4, [0, 0, 2, 2], [1, 1, 2, 0], symmetrise=False)
    assert list(d.offsets) == [0, 1, 1, 3, 3]
    assert list(d.adj) == [1, 0, 2]  # self-loop kept, duplicate removed


def test_csr_from_edges_rejects_out_of_range():
    with pytest.raises(pb.PbbsError):
        pb.from_edges(3, [0, 5], [1, 1])


@pytest.mark.parametrize("d", [1, 2, 3, 6])
def test_grid3d_is_the_symmetrised_torus(d):
    """Same graph as gridGraph -j -d 3 -o d^3 (gridGraph.C:53-72): three
    forward torus edges per vertex, symmetrised, deduplicated, sorted."""
    g = pb.grid3d(d)
    i, j, k = np.meshgrid(np.arange(d), np.arange(d), np.arange(d), indexing="ij")
    loc = lambda a, b, c: ((a % d) * d * d + (b % d) * d + (c % d)).ravel()
    l = loc(i, j, k)
    u = np.concatenate([l, l, l])
    v = np.concatenate([loc(i + 1, j, k), loc(i, j + 1, k), loc(i, j, k + 1)])
    keep = u != v
    u, v = u[keep], v[keep]
    key = np.unique(np.concatenate([u * d ** 3 + v, v * d ** 3 + u]))
    assert g.n == d ** 3
    assert np.array_equal(rows_of(g) * g.n + g.adj, key)
    if d >= 3:
        assert np.all(g.degrees() == 6)


def test_rmat_shape_and_determinism():
    g = pb.rmat(12, 16 * 4096, 0.57, 0.19, 0.19, seed=5)
    assert g.n == 4096
    assert_canonical_symmetric(g)
    h = pb.rmat(12, 16 * 4096, 0.57, 0.19, 0.19, seed=5)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.adj, h.adj)
    o = pb.rmat(12, 16 * 4096, 0.57, 0.19, 0.19, seed=6)
    assert not np.array_equal(g.adj[:1000], o.adj[:1000])
    deg = g.degrees()
    # a=0.57 concentrates edges on low ids: vertex 0 is the largest hub
    assert deg[0] == deg.max()
    # before dedup the low half holds (2a+b+c)/(b+c+2d) = 3.2x the high half
    assert deg[: g.n // 2].sum() > 2 * deg[g.n // 2:].sum()


def test_rmat_quadrant_frequencies():
    """First-level quadrant shares of the draws match (a, b, c, d).  Computed
    on the directed draws recovered from the symmetric graph: an undirected
    edge {u,v} with u,v in different halves came from quadrant b or c."""
    a, b, c = 0.45, 0.15, 0.25
    g = pb.rmat(10, 200_000, a, b, c, seed=1)
    r = rows_of(g)
    half = g.n // 2
    lo_lo = np.mean((r < half) & (g.adj < half))
    hi_hi = np.mean((r >= half) & (g.adj >= half))
    # dedup distorts the shares; a low/low pair is still far the most common
    assert lo_lo > hi_hi * 2


def test_rmat_permutation_relabels():
    g = pb.rmat(10, 8 * 1024, 0.57, 0.19, 0.19, seed=2)
    p = pb.rmat(10, 8 * 1024, 0.57, 0.19, 0.19, seed=2, permute_seed=9)
    assert_canonical_symmetric(p)
    assert g.m == p.m
    assert np.array_equal(np.sort(g.degrees()), np.sort(p.degrees()))
    assert not np.array_equal(g.degrees(), p.degrees())


def test_permute_matches_the_generators_relabelling():
    """permute(g, s) applies the permutation rmat's permute_seed=s applies, so
    relabelling the ordered graph equals generating the relabelled one."""
    g = pb.rmat(10, 8 * 1024, 0.57, 0.19, 0.19, seed=2)
    p = pb.rmat(10, 8 * 1024, 0.57, 0.19, 0.19, seed=2, permute_seed=9)
    q = pb.permute(g, 9)
    assert_canonical_symmetric(q)
    assert np.array_equal(q.offsets, p.offsets) and np.array_equal(q.adj, p.adj)
    assert not np.array_equal(pb.permute(g, 10).adj, q.adj)
    t = pb.permute(pb.grid3d(6), 3)
    assert_canonical_symmetric(t)
    assert (t.degrees() == 6).all() and t.m == 6 * 216


def test_permute_rejects_bad_graphs():
    g = pb.grid3d(3)
    with pytest.raises(pb.PbbsError):
        pb.permute(g, 0)
    bad = pb.HostGraph(g.offsets, np.where(g.adj == 5, 27, g.adj).astype(np.int32))
    with pytest.raises(pb.PbbsError):
        pb.permute(bad, 1)


def test_randlocal_distances():
    n, k = 5000, 10
    g = pb.randlocal(n, k, seed=4)
    assert_canonical_symmetric(g)
    r = rows_of(g)
    dist = np.minimum((g.adj - r) % n, (r - g.adj) % n)
    assert dist.min() >= 1
    # floor(n^u) with u ~ U[0,1): about half the draws are shorter than sqrt(n)
    frac_short = np.mean(dist < np.sqrt(n))
    assert 0.35 < frac_short < 0.65
    assert g.m <= 2 * n * k


def test_generator_argument_errors():
    with pytest.raises(pb.PbbsError):
        pb.rmat(0, 10, 0.5, 0.1, 0.1)
    with pytest.raises(pb.PbbsError):
        pb.rmat(8, 10, 0.7, 0.3, 0.3)
    with pytest.raises(pb.PbbsError):
        pb.grid3d(0)
