"""Autotuner for the direction switch and launch shape of the BFS.

The reference fixes its direction-optimising switch at m/20 (go dense) and
n/20 (stay dense) for every graph (common/ligraLight.h:122,130).  Here those
thresholds, the batched pull threshold and the level-chunk length are
measured per graph class and kept in a small JSON tuning table:

    {"version": 1,
     "entries": {"skewed-s24": {"params": {"td_to_bu_div": 20, ...},
                                "ms": 0.53, "default_ms": 0.54, "graph": {...}}, ...}}

A graph's class is its degree shape ("skewed" when the largest degree is more
than 16 times the mean degree of the non-isolated vertices, else "uniform"),
its scale (log2 n, rounded) and its mean degree (log2 of the mean degree of
the non-isolated vertices, rounded): "uniform-s23-d3" is the 216^3 torus,
"uniform-s23-d4" a random-local graph of 10M vertices and degree ~20.  The search is a coordinate descent over a
small grid per parameter, starting from the library defaults, so the default
setting is always among the measured candidates.  Outputs are bit-identical
under every setting (parents and depths depend only on the graph and the
source); the tuner checks that on every candidate it measures and raises
PbbsError if a digest differs from the default's.

Command line (one GPU):

    python -m pbbsbench_amd.autotune --rmat 22 --out tuning.json
    python -m pbbsbench_amd.autotune --graph g.adj --batch --out tuning.json

and in code: ``autotune.apply(ctx, dg, autotune.load_table("tuning.json"))``.
"""
import argparse
import hashlib
import json
import math
import os
import sys

import numpy as np

from ._lib import PBBS_ERR_FORMAT, PBBS_ERR_INVALID, PBBS_ERR_STATE, PbbsError

# Candidate values per parameter (pbbs_tuning fields, include/pbbs_bfs.h).
SINGLE_GRID = {
    "td_to_bu_div": [5, 10, 20, 40, 80],
    "bu_to_td_div": [5, 10, 20, 40, 80, 160],
    "chunk_levels": [4, 16, 64],
    "bu_serial_depth": [32, 64, 128, 256],
    "persist": [0, 1, 2],
}
BATCH_GRID = {
    "batch_pull_div": [3, 4, 6, 10, 14, 20],
}
TABLE_VERSION = 1


def graph_class(info):
    """Class key of a graph from its info dict (n, m, isolated, max_degree)."""
    n = int(info["n"])
    nz = max(n - int(info["isolated"]), 1)
    mean = max(int(info["m"]) / nz, 1.0)
    shape = "skewed" if int(info["max_degree"]) > 16 * mean else "uniform"
    return (f"{shape}-s{int(round(math.log2(max(n, 2))))}"
            f"-d{int(round(math.log2(mean)))}")


def coordinate_descent(grid, defaults, measure, rounds=2, min_gain=0.01):
    """Minimises measure(params) -> (ms, digest) over the grid, one parameter at
    a time, starting from `defaults` (which need not lie on the grid: each
    parameter's default value is added to its candidates).  A candidate
    replaces the best only when it is faster by more than `min_gain` (a
    fraction), so timing noise does not move the result off the defaults.
    Every candidate's digest must equal the default's.  Returns (best_params,
    best_ms, default_ms, trials) with trials = [(params, ms), ...] in measuring
    order."""
    seen = {}
    trials = []
    ref = {}

    def run(p):
        key = tuple(sorted(p.items()))
        if key not in seen:
            ms, digest = measure(dict(p))
            print(f"[autotune] {dict(p)}: {ms:.4f} ms", file=sys.stderr, flush=True)
            if "digest" not in ref:
                ref["digest"] = digest
            elif digest != ref["digest"]:
                raise PbbsError(PBBS_ERR_STATE,
                               f"outputs differ under tuning {p}: the BFS must not depend on it")
            seen[key] = ms
            trials.append((dict(p), ms))
        return seen[key]

    best = dict(defaults)
    best_ms = default_ms = run(best)
    for _ in range(rounds):
        moved = False
        for name, values in grid.items():
            cands = sorted(set(values) | {defaults[name]})
            for v in cands:
                if v == best[name]:
                    continue
                p = dict(best)
                p[name] = v
                ms = run(p)
                if ms < best_ms * (1.0 - min_gain):
                    best, best_ms, moved = p, ms, True
        if not moved:
            break
    return best, best_ms, default_ms, trials


def pick_sources(dg, count, seed=1):
    """`count` seeded non-isolated sources of a resident graph."""
    deg = dg.download().degrees()
    nz = np.flatnonzero(deg > 0)
    if nz.size == 0:
        raise PbbsError(PBBS_ERR_INVALID, "graph has no edges to tune on")
    rng = np.random.default_rng(seed)
    return rng.choice(nz, min(count, nz.size), replace=False).astype(np.int32)


def _digest(*arrays):
    h = hashlib.sha1()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _median_ms(fn, runs):
    fn()   # warm-up: first-run set-up (row records, graphs) is not timed
    return float(np.median([fn() for _ in range(runs)]))


def tune_single(ctx, dg, sources, runs=5, grid=None, rounds=2):
    """Single-source parameters for dg, timed as the mean over `sources` of the
    median device ms of `runs` BFS runs each."""
    grid = dict(SINGLE_GRID if grid is None else grid)
    base = ctx.tuning
    defaults = {k: int(getattr(base, k)) for k in grid}

    def measure(p):
        ctx.set_tuning(**p)
        try:
            ms = [_median_ms(lambda s=s: dg.bfs_device(int(s))["device_ms"], runs)
                  for s in sources]
            outs = []
            for s in sources:
                par, dep, _ = dg.bfs(int(s))
                outs += [par, dep]
            return float(np.mean(ms)), _digest(*outs)
        finally:
            ctx.set_tuning(**defaults)

    return coordinate_descent(grid, defaults, measure, rounds)


def tune_batch(ctx, dg, sources, runs=5, grid=None, rounds=1):
    """Batched-BFS parameters for dg (one traversal over `sources`, at most 64)."""
    grid = dict(BATCH_GRID if grid is None else grid)
    base = ctx.tuning
    defaults = {k: int(getattr(base, k)) for k in grid}

    def measure(p):
        ctx.set_tuning(**p)
        try:
            ms = _median_ms(lambda: dg.bfs_batch_device(sources)["device_ms"], runs)
            par, dep, _ = dg.bfs_batch(sources)
            return ms, _digest(par, dep)
        finally:
            ctx.set_tuning(**defaults)

    return coordinate_descent(grid, defaults, measure, rounds)


def tune(ctx, dg, runs=5, sources=4, batch=False, seed=1, single_grid=None, batch_grid=None):
    """Measures dg's best settings and returns (class_key, entry) for a table."""
    info = dg.info
    key = graph_class(info)
    src = pick_sources(dg, sources, seed)
    params, ms, default_ms, trials = tune_single(ctx, dg, src, runs, single_grid)
    entry = {
        "params": dict(params),
        "ms": round(ms, 4),
        "default_ms": round(default_ms, 4),
        "trials": len(trials),
        "graph": {k: int(info[k]) for k in ("n", "m", "isolated", "max_degree")},
    }
    if batch:
        bsrc = pick_sources(dg, 64, seed)
        before = {k: int(getattr(ctx.tuning, k)) for k in params}
        ctx.set_tuning(**params)   # the batched search runs under the chosen switch
        try:
            bp, bms, bdef, btr = tune_batch(ctx, dg, bsrc, runs, batch_grid)
        finally:
            ctx.set_tuning(**before)
        entry["params"].update(bp)
        entry["batch_ms"] = round(bms, 4)
        entry["batch_default_ms"] = round(bdef, 4)
        entry["trials"] += len(btr)
    return key, entry


def load_table(path):
    with open(path) as f:
        t = json.load(f)
    if t.get("version") != TABLE_VERSION or not isinstance(t.get("entries"), dict):
        raise PbbsError(PBBS_ERR_FORMAT, f"{path}: not a version-{TABLE_VERSION} tuning table")
    return t


def save_table(path, table):
    tmp = f"{path}.tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)


def merge(table, key, entry):
    t = table if table is not None else {"version": TABLE_VERSION, "entries": {}}
    t["entries"][key] = entry
    return t


def apply(ctx, dg, table):
    """Sets ctx's tuning from the table entry of dg's class.  Returns the key
    applied, or None when the table has no entry for the class (the tuning is
    then left as it was)."""
    key = graph_class(dg.info)
    e = table["entries"].get(key)
    if e is None:
        return None
    known = set(SINGLE_GRID) | set(BATCH_GRID)
    bad = set(e["params"]) - known
    if bad:
        raise PbbsError(PBBS_ERR_FORMAT,
                        f"tuning table entry {key} names unknown parameters {sorted(bad)}")
    ctx.set_tuning(**{k: int(v) for k, v in e["params"].items()})
    return key


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pbbsbench_amd.autotune")
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--rmat", type=int, metavar="SCALE",
                   help="RMAT graph of 2^SCALE vertices, 16 * 2^SCALE draws, a,b,c = .57,.19,.19")
    g.add_argument("--grid", type=int, metavar="D", help="D x D x D torus")
    g.add_argument("--randlocal", type=int, metavar="N",
                   help="random-local graph of N vertices, 10 draws each, seed 1")
    g.add_argument("--graph", metavar="FILE", help="AdjacencyGraph text or PBBSCSR1 binary file")
    ap.add_argument("--out", required=True, help="tuning table (JSON); updated in place")
    ap.add_argument("--wide-offsets", action="store_true", help="keep int64 offsets in HBM")
    ap.add_argument("--batch", action="store_true", help="also tune the batched BFS")
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args(argv)

    from .core import Context
    with Context(0) as ctx:
        if args.rmat is not None:
            dg = ctx.gen_rmat(args.rmat, 16 * 2 ** args.rmat, .57, .19, .19, seed=1,
                              wide_offsets=args.wide_offsets)
        elif args.grid is not None:
            dg = ctx.gen_grid3d(args.grid, wide_offsets=args.wide_offsets)
        elif args.randlocal is not None:
            from .core import randlocal
            dg = ctx.upload(randlocal(args.randlocal, 10, seed=1), wide_offsets=args.wide_offsets)
        else:
            dg = ctx.read_graph(args.graph, wide_offsets=args.wide_offsets)
        key, entry = tune(ctx, dg, args.runs, args.sources, args.batch, args.seed)
        table = load_table(args.out) if os.path.exists(args.out) else None
        save_table(args.out, merge(table, key, entry))
        print(json.dumps({"class": key, **entry}))
        dg.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
