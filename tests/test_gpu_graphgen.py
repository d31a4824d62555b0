"""GPU graph construction (pbbs_graph_gen_*): the CSR built in HBM must equal
the host generator's CSR bit for bit (same seed, same arguments), and a BFS
on it must equal the oracle's."""
import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def same_csr(h, d):
    assert d.n == h.n and d.m == h.m, (d.n, h.n, d.m, h.m)
    assert np.array_equal(d.offsets, h.offsets)
    assert np.array_equal(d.adj, h.adj)


@pytest.mark.parametrize("scale,draws,abc,seed,perm,wide", [
    (10, 16 * 2 ** 10, (0.57, 0.19, 0.19), 1, 0, False),
    (12, 16 * 2 ** 12, (0.5, 0.1, 0.1), 7, 0, True),
    (14, 8 * 2 ** 14, (0.57, 0.19, 0.19), 3, 11, False),
    (16, 2 ** 19, (0.5, 0.1, 0.1), 1, 0, False),          # the C1 workload
    (5, 0, (0.57, 0.19, 0.19), 1, 0, False),               # no edges
    (3, 200, (0.9, 0.05, 0.05), 2, 0, True),               # mostly self-loops and duplicates
])
def test_gen_rmat_equals_host(ctx, scale, draws, abc, seed, perm, wide):
    say(f"rmat scale={scale} draws={draws} abc={abc} seed={seed} perm={perm} wide={wide}")
    h = pb.rmat(scale, draws, *abc, seed=seed, permute_seed=perm)
    dg = ctx.gen_rmat(scale, draws, *abc, seed=seed, permute_seed=perm, wide_offsets=wide)
    try:
        same_csr(h, dg.download())
        info = dg.info
        assert info["symmetric"] and info["offsets_i64"] == int(wide)
    finally:
        dg.free()


@pytest.mark.parametrize("d,wide", [(1, False), (2, False), (3, True), (10, False), (17, True)])
def test_gen_grid3d_equals_host(ctx, d, wide):
    say(f"grid3d d={d} wide={wide}")
    h = pb.grid3d(d)
    dg = ctx.gen_grid3d(d, wide_offsets=wide)
    try:
        same_csr(h, dg.download())
    finally:
        dg.free()


def test_bfs_on_generated_graph(ctx):
    h = pb.rmat(13, 16 * 2 ** 13, 0.57, 0.19, 0.19, seed=5)
    dg = ctx.gen_rmat(13, 16 * 2 ** 13, 0.57, 0.19, 0.19, seed=5)
    try:
        src = int(np.flatnonzero(h.degrees() > 0)[3])
        parent, depth, _ = dg.bfs(src)
        wp, wd = oracle.serial_bfs(h.offsets, h.adj, src)
        assert np.array_equal(parent, wp) and np.array_equal(depth, wd)
        P, D, _ = dg.bfs_batch([src, 0, 17])
        assert np.array_equal(P[0], wp) and np.array_equal(D[0], wd)
    finally:
        dg.free()


def test_gen_rejects_bad_arguments(ctx):
    with pytest.raises(pb.PbbsError):
        ctx.gen_rmat(31, 10, 0.5, 0.1, 0.1)
    with pytest.raises(pb.PbbsError):
        ctx.gen_rmat(8, 10, 0.7, 0.3, 0.3)
    with pytest.raises(pb.PbbsError):
        ctx.gen_grid3d(0)
