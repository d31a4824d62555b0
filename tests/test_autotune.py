"""Autotuner search logic and tuning-table format, CPU only (the measuring
function is a synthetic cost here; tests/test_gpu_autotune.py measures)."""
import json

import numpy as np
import pytest

import pbbsbench_amd as pb
from pbbsbench_amd import autotune


def host_info(g):
    deg = g.degrees()
    return {"n": g.n, "m": g.m, "isolated": int((deg == 0).sum()), "max_degree": int(deg.max())}


def test_graph_class():
    assert autotune.graph_class(host_info(pb.rmat(12, 16 * 4096, .57, .19, .19, seed=1))) == \
        "skewed-s12-d5"
    assert autotune.graph_class(host_info(pb.grid3d(16))) == "uniform-s12-d3"
    assert autotune.graph_class(host_info(pb.randlocal(100000, 10, seed=1))) == "uniform-s17-d4"
    # rounding of the scale and the 16x-mean cut
    assert autotune.graph_class({"n": 3000, "m": 30000, "isolated": 0, "max_degree": 160}) == \
        "uniform-s12-d3"
    assert autotune.graph_class({"n": 3000, "m": 30000, "isolated": 0, "max_degree": 161}) == \
        "skewed-s12-d3"
    # the 216^3 torus and a random-local graph of 10M vertices share a scale
    # but not a class (the mean degree tells them apart)
    assert autotune.graph_class({"n": 10077696, "m": 60466176, "isolated": 0,
                                 "max_degree": 6}) == "uniform-s23-d3"
    assert autotune.graph_class({"n": 10_000_000, "m": 196714352, "isolated": 0,
                                 "max_degree": 40}) == "uniform-s23-d4"


def separable_cost(opt):
    """A cost with its minimum at `opt`, one term per parameter."""
    def measure(p):
        ms = 1.0 + sum(abs(np.log2(p[k] / opt[k])) * 0.1 for k in opt)
        return ms, "same"
    return measure


def test_descent_finds_the_grid_minimum():
    grid = {"a": [1, 2, 4, 8, 16], "b": [10, 20, 40, 80]}
    best, ms, dms, trials = autotune.coordinate_descent(grid, {"a": 4, "b": 20},
                                                        separable_cost({"a": 16, "b": 80}))
    assert best == {"a": 16, "b": 80}
    assert ms == pytest.approx(1.0) and dms > ms
    assert trials[0] == ({"a": 4, "b": 20}, dms)          # the default is measured first
    keys = [tuple(sorted(p.items())) for p, _ in trials]
    assert len(keys) == len(set(keys))                       # no setting measured twice


def test_descent_adds_the_default_to_the_candidates():
    grid = {"a": [1, 2]}
    best, ms, dms, trials = autotune.coordinate_descent(grid, {"a": 3},
                                                        separable_cost({"a": 3}))
    assert best == {"a": 3} and ms == dms
    assert sorted(p["a"] for p, _ in trials) == [1, 2, 3]


def test_descent_ignores_gains_below_the_threshold():
    def measure(p):
        return (0.995 if p["a"] == 2 else 1.0), "same"
    best, _, _, _ = autotune.coordinate_descent({"a": [1, 2]}, {"a": 1}, measure)
    assert best == {"a": 1}
    best, _, _, _ = autotune.coordinate_descent({"a": [1, 2]}, {"a": 1}, measure, min_gain=0.001)
    assert best == {"a": 2}


def test_descent_rejects_settings_that_change_outputs():
    def measure(p):
        return 1.0 / p["a"], "base" if p["a"] == 1 else "other"
    with pytest.raises(pb.PbbsError, match="outputs differ"):
        autotune.coordinate_descent({"a": [1, 2]}, {"a": 1}, measure)


def test_table_round_trip(tmp_path):
    p = tmp_path / "t.json"
    e1 = {"params": {"td_to_bu_div": 40, "bu_to_td_div": 20, "chunk_levels": 16}, "ms": 0.5,
          "default_ms": 0.52, "trials": 9, "graph": {"n": 1, "m": 0, "isolated": 1,
                                                     "max_degree": 0}}
    t = autotune.merge(None, "skewed-s24", e1)
    autotune.save_table(str(p), t)
    t2 = autotune.load_table(str(p))
    assert t2 == t and t2["version"] == autotune.TABLE_VERSION
    e2 = dict(e1, ms=0.4)
    t3 = autotune.merge(t2, "uniform-s20", e2)
    assert sorted(t3["entries"]) == ["skewed-s24", "uniform-s20"]
    p.write_text(json.dumps({"version": 99, "entries": {}}))
    with pytest.raises(pb.PbbsError, match="tuning table"):
        autotune.load_table(str(p))


def test_grids_name_tuning_fields():
    from pbbsbench_amd import _lib
    fields = {f for f, _ in _lib.Tuning._fields_}
    assert set(autotune.SINGLE_GRID) <= fields and set(autotune.BATCH_GRID) <= fields
