"""The benchmark configurations at full size, checked through properties
that hold at any size: the GPU checker's BFSCheck rule, the max-parent rule
and depth == level for every output row; the torus against its closed form.
Graphs are built on the GPU (bit-identical to the host generators, see
test_gpu_graphgen.py), so each case takes seconds."""
import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def nonisolated_sources(dg, k, seed):
    g = dg.download()
    rng = np.random.default_rng(seed)
    return rng.choice(np.flatnonzero(g.degrees() > 0), k, replace=False).astype(np.int32), g


def checked(dg, ctx, s, stats, batched_row=None):
    pp, dp = ctx.outputs(batched=batched_row is not None)
    if batched_row is not None:
        pp += 4 * batched_row * dg.n
        dp += 4 * batched_row * dg.n
    r = dg.check_device(int(s), pp, dp)
    assert r["valid"] == 1 and r["cycle"] == 0, r
    assert r["max_rule_bad"] == 0 and r["depth_bad"] == 0, r
    return r


def test_c4_rmat24_int64(ctx):
    say("C4 rmat24")
    dg = ctx.gen_rmat(24, 2 ** 28, 0.57, 0.19, 0.19, seed=1, wide_offsets=True)
    try:
        src, g = nonisolated_sources(dg, 3, 1)
        for s in src:
            st = dg.bfs_device(int(s))
            r = checked(dg, ctx, s, st)
            assert r["reached"] == st["reached"] and r["max_level"] + 1 == st["levels"]
    finally:
        dg.free()


def test_c5a_rmat22_batch64(ctx):
    say("C5a rmat22 x 64")
    dg = ctx.gen_rmat(22, 2 ** 26, 0.57, 0.19, 0.19, seed=1)
    try:
        src, g = nonisolated_sources(dg, 64, 1)
        st = dg.bfs_batch_device(src)
        total = 0
        for i, s in enumerate(src):
            total += checked(dg, ctx, s, st, batched_row=i)["reached"]
        assert total == st["reached"]
    finally:
        dg.free()


def test_c3_torus216_closed_form(ctx):
    say("C3 torus 216")
    dg = ctx.gen_grid3d(216)
    try:
        parent, depth, st = dg.bfs(0)
        assert st["levels"] == 325
        wp, wd = oracle.torus_closed_form(216)
        assert np.array_equal(depth, wd)
        assert np.array_equal(parent, wp)
    finally:
        dg.free()


def test_c2_randlocal10m(ctx):
    say("C2 randlocal 10M")
    g = pb.randlocal(10_000_000, 10, seed=1)
    dg = ctx.upload(g, wide_offsets=True)
    try:
        st = dg.bfs_device(0)
        checked(dg, ctx, 0, st)
    finally:
        dg.free()


@pytest.mark.parametrize("grid", [0, 256])
def test_persistent_launch_full_size(grid):
    """Every level in one persistent launch (grid barrier between levels) on
    the C3 torus and the C4 RMAT graph: the torus equals its closed form, the
    RMAT outputs pass the checker and equal the one-level-per-launch run."""
    with pb.Context(0) as c:
        c.set_tuning(persist=1, grid_blocks=grid)
        say(f"persistent, grid {grid or 'auto'}: C3 torus 216")
        dg = c.gen_grid3d(216)
        try:
            parent, depth, st = dg.bfs(0)
            assert st["levels"] == 325 and st["launches"] == 1
            wp, wd = oracle.torus_closed_form(216)
            assert np.array_equal(depth, wd) and np.array_equal(parent, wp)
        finally:
            dg.free()
        say(f"persistent, grid {grid or 'auto'}: C4 rmat24")
        dg = c.gen_rmat(24, 2 ** 28, 0.57, 0.19, 0.19, seed=1, wide_offsets=True)
        try:
            src, _ = nonisolated_sources(dg, 2, 2)
            for s in src:
                p1, d1, st = dg.bfs(int(s))
                assert st["launches"] == 1
                r = dg.check(int(s), p1, d1)
                assert r["valid"] == 1 and r["max_rule_bad"] == 0 and r["depth_bad"] == 0, r
                c.set_tuning(persist=0)
                p0, d0, _ = dg.bfs(int(s))
                c.set_tuning(persist=1)
                assert np.array_equal(p1, p0) and np.array_equal(d1, d0)
        finally:
            dg.free()
