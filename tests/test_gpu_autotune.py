"""The autotuner on the GPU: it measures a small grid on real graphs, keeps
outputs bit-identical across every setting it tries (it raises otherwise), and
the table it writes sets the context's tuning for the graph's class."""
import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from pbbsbench_amd import autotune

pytestmark = pytest.mark.gpu

SMALL_SINGLE = {"td_to_bu_div": [5, 20, 80], "bu_to_td_div": [10, 40], "chunk_levels": [4, 16]}
SMALL_BATCH = {"batch_pull_div": [4, 20]}


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("kind", ["rmat14", "torus24"])
def test_tune_apply_and_parity(ctx, tmp_path, kind):
    say(f"autotune {kind}")
    saved = {f: int(getattr(ctx.tuning, f)) for f, _ in ctx.tuning._fields_}
    dg = ctx.gen_rmat(14, 16 * 2 ** 14, .57, .19, .19, seed=3) if kind == "rmat14" \
        else ctx.gen_grid3d(24)
    try:
        key, entry = autotune.tune(ctx, dg, runs=2, sources=2, batch=True, seed=5,
                                   single_grid=SMALL_SINGLE, batch_grid=SMALL_BATCH)
        assert key == autotune.graph_class(dg.info)
        assert key.startswith("skewed-s14-" if kind == "rmat14" else "uniform-s14-")
        p = entry["params"]
        for k, vals in list(SMALL_SINGLE.items()) + list(SMALL_BATCH.items()):
            assert p[k] in set(vals) | {saved[k]}
        assert entry["ms"] <= entry["default_ms"] and entry["batch_ms"] <= entry["batch_default_ms"]
        # the search left the context's tuning as it found it
        assert {f: int(getattr(ctx.tuning, f)) for f, _ in ctx.tuning._fields_} == saved

        path = str(tmp_path / "tuning.json")
        autotune.save_table(path, autotune.merge(None, key, entry))
        assert autotune.apply(ctx, dg, autotune.load_table(path)) == key
        for k, v in p.items():
            assert int(getattr(ctx.tuning, k)) == v
        # outputs under the applied tuning equal the oracle's
        g = dg.download()
        src = autotune.pick_sources(dg, 3, seed=9)
        for s in src:
            par, dep, _ = dg.bfs(int(s))
            wp, wd = oracle.serial_bfs(g.offsets, g.adj, int(s))
            assert np.array_equal(par, wp) and np.array_equal(dep, wd), f"source {s}"
        bp, bd, _ = dg.bfs_batch(src)
        for i, s in enumerate(src):
            wp, wd = oracle.serial_bfs(g.offsets, g.adj, int(s))
            assert np.array_equal(bp[i], wp) and np.array_equal(bd[i], wd), f"batch row {i}"
        # a class the table does not hold leaves the tuning alone
        other = ctx.gen_grid3d(8) if kind == "rmat14" else ctx.gen_rmat(10, 16 * 1024, .57, .19,
                                                                        .19, seed=1)
        try:
            before = {f: int(getattr(ctx.tuning, f)) for f, _ in ctx.tuning._fields_}
            assert autotune.apply(ctx, other, autotune.load_table(path)) is None
            assert {f: int(getattr(ctx.tuning, f)) for f, _ in ctx.tuning._fields_} == before
        finally:
            other.free()
    finally:
        ctx.set_tuning(**saved)
        dg.free()
