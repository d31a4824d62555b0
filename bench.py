#!/usr/bin/env python3
"""Benchmark of the deterministic BFS on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmat24]

The default run measures the three parts of the metric and prints ONE JSON
line:

* single-source BFS (the top-level line).  A step is one BFS over the whole
  graph (init + every level), with the CSR already resident in HBM and the
  parent/depth outputs written to HBM.  The default workload is the RMAT graph
  the target is quoted on: 2^24 vertices, 2^28 directed draws with
  (a,b,c,d) = (.57,.19,.19,.05), self-loops and duplicates removed,
  symmetrised, int64 offsets, source = a seeded random non-isolated vertex.
  `value` is GTEPS = m_r / t with m_r the directed CSR entries of the reached
  vertices and t the wall-clock time per step over the K timed steps.
  `roofline` prices one BFS launch sequence against HBM: algorithmic bytes
  B = S_o*(n+1) + 4*m_r + 8*n over the hipEvent time of the sequence,
  measured on the context's stream over the timed region.
* `batched`: 64 seeded sources in one traversal on RMAT 2^22 against 64
  sequential single-source runs of the same sources (ratio by wall time and
  by device time), bit-exact against those runs and a sample of the oracle.
* `serving`: 8 resident RMAT graphs of 2^16..2^22 vertices, a burst and then
  Poisson arrivals; p50/p99 latency (arrival -> parents in host memory),
  throughput and mean batch; a seeded sample compared with the oracle.

`cpu_baseline` times the oracle's C restatement of deterministicBFS on the
host cores with the protocol of common/time_loop.h:4-18 (>= 1 s of untimed
warm-up, then timed rounds; min and median reported), since the reference's
parlaylib build cannot be compiled here.  The same run checks the GPU output
bit for bit.  Per-graph search data (row-top records, share bounds, degree
ranks) are built before the timed region; their time and a cold first call
are reported in `config`.
"""
import argparse
import json
import os
import sys
import time
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("GTEPS + % HBM roofline (1 GPU); batched-64-source vs sequential GTEPS; "
          "p50/p99 latency")
HBM_PEAK_GBS = 8000.0
TUNING_TABLE = os.path.join(ROOT, "pbbsbench_amd", "tuning_mi355x.json")

# GPU builders (ctx, wide) -> DeviceGraph; the same graphs as the host builders,
# bit for bit (tests/test_gpu_graphgen.py), built in HBM in milliseconds.
GPU_BUILDERS = {
    "rmat24": lambda ctx, w: ctx.gen_rmat(24, 2 ** 28, 0.57, 0.19, 0.19, seed=1, wide_offsets=w),
    "rmat16": lambda ctx, w: ctx.gen_rmat(16, 2 ** 19, 0.5, 0.1, 0.1, seed=1, wide_offsets=w),
    "torus216": lambda ctx, w: ctx.gen_grid3d(216, wide_offsets=w),
    "rmat22": lambda ctx, w: ctx.gen_rmat(22, 2 ** 26, 0.57, 0.19, 0.19, seed=1, wide_offsets=w),
}

WORKLOADS = {
    # name: (host builder, wide offsets, description)
    "rmat24": (lambda: __import__("pbbsbench_amd").rmat(24, 2 ** 28, 0.57, 0.19, 0.19, seed=1),
               True, "RMAT 2^24 vertices, 2^28 draws, a=.57 b=.19 c=.19, symmetrised, int64 offsets"),
    "rmat16": (lambda: __import__("pbbsbench_amd").rmat(16, 2 ** 19, 0.5, 0.1, 0.1, seed=1),
               False, "RMAT 2^16 vertices, 2^19 draws, a=.5 b=.1 c=.1, symmetrised, uint32 offsets"),
    "randlocal10m": (lambda: __import__("pbbsbench_amd").randlocal(10_000_000, 10, seed=1),
                     True, "random-local 10M vertices, 10 draws each, symmetrised, int64 offsets"),
    "torus216": (lambda: __import__("pbbsbench_amd").grid3d(216),
                 False, "3D torus 216^3, degree 6, uint32 offsets"),
    "rmat22": (lambda: __import__("pbbsbench_amd").rmat(22, 2 ** 26, 0.57, 0.19, 0.19, seed=1),
               True, "RMAT 2^22 vertices, 2^26 draws, a=.57 b=.19 c=.19, symmetrised, int64 offsets"),
}
SOURCE_ZERO = {"rmat16", "randlocal10m", "torus216"}
SERVE_GRAPHS = [(16, 1), (17, 1), (18, 1), (19, 1), (20, 1), (21, 1), (22, 1), (22, 2)]


def pick_source(g, name, seed):
    if name in SOURCE_ZERO:
        return 0
    deg = g.degrees()
    rng = np.random.default_rng(seed)
    while True:
        s = int(rng.integers(0, g.n))
        if deg[s] > 0:
            return s


def say(*a):
    """Progress of the run on stderr (the JSON line alone goes to stdout)."""
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)


def apply_tuning(ctx, dg, use_table):
    """Applies the committed tuning table's entry for dg's class (if any) and
    returns what config records about it."""
    from pbbsbench_amd import autotune
    if not use_table or not os.path.exists(TUNING_TABLE):
        return {"table": None, "class": autotune.graph_class(dg.info), "applied": False}
    key = autotune.apply(ctx, dg, autotune.load_table(TUNING_TABLE))
    t = ctx.tuning
    return {"table": os.path.relpath(TUNING_TABLE, ROOT), "class": autotune.graph_class(dg.info),
            "applied": key is not None,
            "params": {k: int(getattr(t, k)) for k, _ in t._fields_}}


def cpu_baseline(g, sources, budget_s, label):
    """The oracle's deterministicBFS restatement timed with the time_loop
    protocol (common/time_loop.h:4-18): untimed runs for at least one second,
    then up to 5 timed rounds while the budget lasts.  `sources` are cycled
    (one source for the single-source line).  Returns (record, outputs of the
    first source)."""
    import oracle  # the checker and the CPU baseline; never the measured product
    threads = host_threads()
    deg = g.degrees()
    t0 = time.perf_counter()
    warm = 0
    first = None
    while time.perf_counter() - t0 < 1.0:
        p, _ = oracle.det_bfs(g.offsets, g.adj, int(sources[warm % len(sources)]), threads=threads)
        if warm == 0:
            first = p
        warm += 1
    times, edges = [], []
    t1 = time.perf_counter()
    while len(times) < 5 and (not times or time.perf_counter() - t1 + times[-1] < budget_s):
        s = int(sources[len(times) % len(sources)])
        t2 = time.perf_counter()
        p, _ = oracle.det_bfs(g.offsets, g.adj, s, threads=threads)
        times.append(time.perf_counter() - t2)
        edges.append(int(deg[p != -1].sum()))
    tmed = float(np.median(times))
    rate = float(np.median([e / t for e, t in zip(edges, times)]))
    rec = {
        "value": round(rate / 1e9, 3),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "min_s": round(float(min(times)), 4),
        "median_s": round(tmed, 4),
        "sample": (f"{label}: {warm} untimed warm-up runs (>= 1 s), then {len(times)} timed "
                   f"rounds (min {min(times):.3f} s, median {tmed:.3f} s) of the oracle's C "
                   f"restatement of deterministicBFS (parlaylib unavailable), {threads} host "
                   f"threads; CPU: {cpu_info()}"),
    }
    return rec, first


def traffic_from_profile(name, **match):
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        tr = json.load(f)
    if all(tr.get(k) == v for k, v in match.items()):
        return tr.get("bytes_per_bfs", tr.get("bytes_per_batch"))
    return None


# ---------------------------------------------------------------- single source
def time_any_parent(args, dg, src, deg):
    """The same BFS under the backForwardBFS contract (PBBS_BFS_ANY_PARENT,
    backForwardBFS/BFS.C:38-61): device time on the bench source and on the
    spread sources, and the GPU checker's verdict on the bench source (a valid
    BFS tree with exact depths).  Not part of `value`."""
    for _ in range(max(args.warmup, 1)):
        dg.bfs_device(src, any_parent=True)
    dev, t0 = [], time.perf_counter()
    for _ in range(args.steps):
        st = dg.bfs_device(src, any_parent=True)
        dev.append(st["device_ms"])
    wall_ms = (time.perf_counter() - t0) / args.steps * 1e3
    p, d, _ = dg.bfs(src, any_parent=True)
    chk = dg.check(src, p, d)
    out = {"contract": "any parent one level up, exact depths (backForwardBFS/BFS.C:38-61)",
           "event_ms_per_bfs": round(float(np.mean(dev)), 4),
           "ms_per_step": round(wall_ms, 4),
           "gteps": round(st["edges_reached"] / (wall_ms / 1e3) / 1e9, 2),
           "check": {"valid": chk["valid"], "depth_bad": chk["depth_bad"], "bad": chk["bad"]}}
    if args.spread > 0:
        rng = np.random.default_rng(args.seed + 1000)
        others = rng.choice(np.flatnonzero(deg > 0), args.spread, replace=False)
        sms, sb = [], []
        So = 8 if dg.info["offsets_i64"] else 4
        for s in others:
            dg.bfs_device(int(s), any_parent=True)
            rs = [dg.bfs_device(int(s), any_parent=True) for _ in range(3)]
            sms.append(float(np.median([r["device_ms"] for r in rs])))
            sb.append(So * (dg.n + 1) + 4 * rs[-1]["edges_reached"] + 8 * dg.n)
        out["spread_event_ms_mean"] = round(float(np.mean(sms)), 4)
        out["spread_event_ms_max"] = round(float(np.max(sms)), 4)
        out["spread_roofline_frac_aggregate"] = round(
            float(sum(sb)) / (float(sum(sms)) / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    say(f"single: any-parent {out['event_ms_per_bfs']} ms per BFS (device)")
    return out


def run_single(args):
    import pbbsbench_amd as pb

    build, wide, desc = WORKLOADS[args.workload]
    ctx = pb.Context(0)
    tune = dict(args.tune)
    if args.td_to_bu_div is not None:
        tune["td_to_bu_div"] = args.td_to_bu_div
    if args.bu_to_td_div is not None:
        tune["bu_to_td_div"] = args.bu_to_td_div
    t0 = time.time()
    on_gpu = args.workload in GPU_BUILDERS and not args.host_gen
    if on_gpu:
        dg = GPU_BUILDERS[args.workload](ctx, wide)   # built in HBM
        gen_s = time.time() - t0
        g = dg.download()   # host copy for source choice, CPU baseline and check
        upload_s = 0.0
    else:
        g = build()
        gen_s = time.time() - t0
        t0 = time.time()
        dg = ctx.upload(g, wide_offsets=wide)
        upload_s = time.time() - t0
    tuning = apply_tuning(ctx, dg, not tune and not args.no_tuning_table)
    if tune:
        ctx.set_tuning(**tune)
        tuning = {"table": None, "applied": False, "params": dict(tune)}
    src = pick_source(g, args.workload, args.seed)
    deg = g.degrees()
    say(f"single: {args.workload} n={g.n} m={g.m} built in {gen_s:.1f} s, source {src}")

    # A cold call: the first BFS on a graph builds its search data (row-top
    # records, share bounds) and captures the launch graphs.  Then a second
    # copy of the graph times pbbs_graph_prepare alone.
    t0 = time.perf_counter()
    dg.bfs_device(src)
    cold_ms = (time.perf_counter() - t0) * 1e3
    prep_ms = None
    if on_gpu:
        dg2 = GPU_BUILDERS[args.workload](ctx, wide)
        prep_ms = dg2.prepare(single=True, batch=False)
        dg2.free()

    for _ in range(args.warmup):
        dg.bfs_device(src)
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = dg.bfs_device(src)  # returns once outputs are in HBM
        dev_ms.append(st["device_ms"])
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    m_r = st["edges_reached"]
    n = g.n
    So = 8 if wide else 4
    B = So * (n + 1) + 4 * m_r + 8 * n
    event_ms = float(np.mean(dev_ms))
    gteps = m_r / (ms_per_step / 1e3) / 1e9

    # Device time depends on the source: a source whose big level is reached
    # through a frontier of a few low-id hubs costs more than the bench source.
    # Other seeded sources are timed after the timed region.
    say(f"single: {args.steps} timed steps, {ms_per_step:.3f} ms per step (wall)")
    spread = None
    if args.spread > 0:
        rng = np.random.default_rng(args.seed + 1000)
        others = rng.choice(np.flatnonzero(deg > 0), args.spread, replace=False)
        sms, smr = [], []
        for s in others:
            dg.bfs_device(int(s))   # warm-up: the first chunk adapts to the source's depth
            rs = [dg.bfs_device(int(s)) for _ in range(3)]
            sms.append(float(np.median([r["device_ms"] for r in rs])))
            smr.append(rs[-1]["edges_reached"])
        sms_a = np.array(sms)
        So_b = [So * (n + 1) + 4 * e + 8 * n for e in smr]
        spread = {
            "sources": int(args.spread), "seed": args.seed + 1000,
            "event_ms_min": round(float(sms_a.min()), 4),
            "event_ms_median": round(float(np.median(sms_a)), 4),
            "event_ms_mean": round(float(sms_a.mean()), 4),
            "event_ms_max": round(float(sms_a.max()), 4),
            "gteps_device_aggregate": round(float(sum(smr)) / (float(sms_a.sum()) / 1e3) / 1e9, 2),
            "roofline_frac_aggregate": round(float(sum(So_b)) / (float(sms_a.sum()) / 1e3) / 1e9
                                             / HBM_PEAK_GBS, 4),
        }

    # outputs to host once (D2H reported, not timed) and bit-exact check
    say("single: source spread done" if spread else "single: no source spread")
    any_parent = None if args.no_any_parent else time_any_parent(args, dg, src, deg)
    parent, depth, st_host = dg.bfs(src)
    result = {
        "metric": METRIC,
        "value": round(gteps, 2),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generator in libpbbs_bfs)",
        "config": {
            "workload": args.workload,
            "description": desc,
            "n": n, "m": g.m, "source": src, "offsets": "int64" if wide else "uint32",
            "reached": st["reached"], "m_r": m_r, "levels": st["levels"],
            "td_levels": st["td_levels"], "bu_levels": st["bu_levels"],
            "event_ms_per_bfs": round(event_ms, 4),
            "gteps_graph500": round(gteps / 2, 2),
            "d2h_ms": round(st_host["copy_ms"], 3),
            "cold_first_bfs_ms": round(cold_ms, 3),
            "preprocess_ms": None if prep_ms is None else round(prep_ms, 3),
            "preprocess_note": ("per-graph search data built once and kept (row-top records, "
                                "bottom-up share bounds; pbbs_graph_prepare), outside the timed "
                                "region; cold_first_bfs_ms is the first BFS on a fresh graph, "
                                "which builds them and captures the launch graphs"),
            "gen_s": round(gen_s, 3), "gen": "gpu" if on_gpu else "host",
            "upload_s": round(upload_s, 2),
            "tuning": tuning,
            "parallelism": "single GPU",
            "source_spread": spread,
            "any_parent": any_parent,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("one BFS: bfs_init_kernel + one bfs_persist_kernel launch running every "
                       "level" if st["launches"] == 1 else
                       "one BFS launch sequence (bfs_init_kernel + bfs_level_kernel per level)"),
            "launches_per_bfs": st["launches"],
            "achieved": round(B / (event_ms / 1e3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(B / (event_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile(args.workload, n=n, m=g.m, source=src),
            "algorithmic_bytes": int(B),
        },
    }
    exact = True
    if args.cpu_seconds > 0:
        say(f"single: CPU baseline (budget {args.cpu_seconds} s after a 1 s warm-up)")
        rec, ref = cpu_baseline(g, [src], args.cpu_seconds,
                                f"full BFS runs from the bench source on {args.workload}")
        exact = bool(np.array_equal(ref, parent))
        import oracle
        exact = exact and bool(np.array_equal(oracle.level_number(ref), depth))
        result["cpu_baseline"] = rec
        result["parity"] = "bit-exact parent and depth vs oracle" if exact else "MISMATCH"
    dg.free()
    ctx.close()
    return result, exact


# ---------------------------------------------------------------------- batched
def run_batch(args, nested):
    """64 sources in one batched traversal vs 64 sequential single-source runs
    on RMAT 2^22 (2^26 draws, a=.57 b=.19 c=.19), uint32 offsets.  A step is
    one batched BFS of all 64 sources with outputs [64][n] written to HBM."""
    import oracle
    import pbbsbench_amd as pb
    ctx = pb.Context(0)
    t0 = time.time()
    if args.host_gen:
        g = pb.rmat(22, 2 ** 26, 0.57, 0.19, 0.19, seed=1)
        gen_s = time.time() - t0
        dg = ctx.upload(g)
    else:
        dg = GPU_BUILDERS["rmat22"](ctx, False)
        gen_s = time.time() - t0
        g = dg.download()
    tuning = apply_tuning(ctx, dg, not args.no_tuning_table and not args.tune)
    if args.tune:
        ctx.set_tuning(**dict(args.tune))
        tuning = {"table": None, "fields": dict(args.tune), "applied": True}
    prep_ms = dg.prepare(single=True, batch=True)
    deg = g.degrees()
    rng = np.random.default_rng(args.seed)
    src = rng.choice(np.flatnonzero(deg > 0), 64, replace=False).astype(np.int32)
    t1 = time.perf_counter()
    dg.bfs_batch_device(src)
    first_ms = (time.perf_counter() - t1) * 1e3
    for _ in range(max(args.warmup, 1)):
        dg.bfs_batch_device(src)
        for s in src:
            dg.bfs_device(int(s))
    say(f"batched: rmat22 n={g.n} m={g.m}, preprocessing {prep_ms:.1f} ms, warm-up done")
    # sequential: 64 single-source runs, outputs in HBM
    seq_ms, seq_edges = 0.0, 0
    t0 = time.perf_counter()
    for s in src:
        st = dg.bfs_device(int(s))
        seq_ms += st["device_ms"]
        seq_edges += st["edges_reached"]
    seq_wall = (time.perf_counter() - t0) * 1e3
    # batched
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = dg.bfs_batch_device(src)
        dev_ms.append(st["device_ms"])
    wall = (time.perf_counter() - t0) / args.steps * 1e3
    pairs_edges = st["edges_reached"]
    exact = pairs_edges == seq_edges
    say(f"batched: {wall:.3f} ms per batch (wall), 64 sequential runs {seq_wall:.3f} ms")
    # parity: every batched row equals its sequential run, bit for bit, and a
    # seeded sample of rows equals the oracle
    P, D, _ = dg.bfs_batch(src)
    for i, s in enumerate(src):
        p, d, _ = dg.bfs(int(s))
        exact &= bool(np.array_equal(P[i], p)) and bool(np.array_equal(D[i], d))
    sample = rng.choice(64, 4, replace=False)
    for i in sample:
        ref, _ = oracle.det_bfs(g.offsets, g.adj, int(src[i]), threads=host_threads())
        exact &= bool(np.array_equal(ref, P[i])) and bool(
            np.array_equal(oracle.level_number(ref), D[i]))
    reached_any = (P != -1).any(axis=0)
    m_union = int(deg[reached_any].sum())
    n = g.n
    B = 4 * (n + 1) + 4 * m_union + 8 * 64 * n
    ev = float(np.mean(dev_ms))
    res = {
        "workload": "rmat22x64",
        "description": "RMAT 2^22 vertices, 2^26 draws, a=.57 b=.19 c=.19, symmetrised, "
                       "uint32 offsets; 64 seeded non-isolated sources in one traversal",
        "n": n, "m": g.m, "sources_seed": args.seed, "steps": args.steps,
        "value_gteps": round(pairs_edges / (wall / 1e3) / 1e9, 2),
        "ms_per_batch": round(wall, 3),
        "event_ms_per_batch": round(ev, 3),
        "pairs_reached": st["reached"], "sum_m_r": pairs_edges,
        "levels": st["levels"], "push_levels": st["td_levels"], "pull_levels": st["bu_levels"],
        "sequential_wall_ms_64": round(seq_wall, 3),
        "sequential_event_ms_64": round(seq_ms, 3),
        "sequential_gteps": round(seq_edges / (seq_wall / 1e3) / 1e9, 2),
        "batch_vs_sequential": round(seq_wall / wall, 3),
        "batch_vs_sequential_device": round(seq_ms / ev, 3),
        "preprocess_ms": round(prep_ms, 3),
        "first_batch_ms": round(first_ms, 3),
        "gen_s": round(gen_s, 3), "gen": "host" if args.host_gen else "gpu",
        "tuning": tuning,
        "roofline": {
            "bound": "hbm",
            "kernel": "one batched BFS launch sequence (mbfs clear/seed, mbfs_level_kernel per "
                      "level, mbfs_depth_kernel)",
            "achieved": round(B / (ev / 1e3) / 1e9, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(B / (ev / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile("rmat22x64", n=n, m=g.m, sources_seed=args.seed),
            "algorithmic_bytes": int(B),
        },
        "parity": ("bit-exact vs the 64 sequential runs and vs the oracle on 4 seeded rows"
                   if exact else "MISMATCH"),
    }
    if not nested and args.cpu_seconds > 0:
        rec, _ = cpu_baseline(g, src, args.cpu_seconds,
                              "single-source BFS runs from the batch's sources, one after another")
        res["cpu_baseline"] = rec
    dg.free()
    ctx.close()
    return res, exact


# ---------------------------------------------------------------------- serving
def run_serve(args, nested):
    """Query stream over 8 resident RMAT graphs (scales 16..22 plus a second
    2^22 graph; 16 draws per vertex, a=.57 b=.19 c=.19).  Each query picks a
    graph and a non-isolated source at random and asks for parents in a host
    buffer.  First a burst of queries submitted at once gives the peak
    throughput; then Poisson arrivals at --qps (default: half that peak) give
    the latency percentiles, arrival -> parents in the caller's buffer.  A
    seeded sample of the Poisson queries is compared with the oracle.  Output
    buffers come from per-graph pools allocated and touched before timing, as
    a serving client reuses its buffers."""
    import threading

    import oracle
    import pbbsbench_amd as pb
    rng = np.random.default_rng(args.seed)
    graphs, nz = [], []
    t0 = time.time()
    with pb.Context(0) as gctx:   # the same graphs as pb.rmat(...), built on the GPU
        for scale, seed in SERVE_GRAPHS:
            dg = gctx.gen_rmat(scale, 16 * 2 ** scale, 0.57, 0.19, 0.19, seed=seed)
            g = dg.download()
            dg.free()
            graphs.append(g)
            nz.append(np.flatnonzero(g.degrees() > 0))
    gen_s = time.time() - t0
    G = len(graphs)

    def draw(q):
        gsel = rng.integers(0, G, q)
        return gsel, [int(nz[gi][rng.integers(0, nz[gi].size)]) for gi in gsel]

    burst = min(args.queries, 512)
    Q = args.queries
    pools = [[] for _ in range(G)]
    lock = threading.Lock()

    def take(gi):
        with lock:
            if pools[gi]:
                return pools[gi].pop()
        return np.full(graphs[gi].n, -2, dtype=np.int32)

    def give(gi, buf):
        with lock:
            pools[gi].append(buf)

    srv = pb.Server(0)
    ids = [srv.add_graph(g) for g in graphs]
    srv.drain([srv.submit(ids[i], int(nz[i][0]), depth=False) for i in range(G)])

    say(f"serving: {G} graphs resident")
    # burst: peak throughput
    bsel, bsrc = draw(burst)
    bufs = [take(gi) for gi in bsel]
    t0 = time.perf_counter()
    tks = [srv.submit(ids[bsel[q]], bsrc[q], out=(bufs[q], None)) for q in range(burst)]
    got = srv.drain(tks)
    peak = burst / (time.perf_counter() - t0)
    burst_batch = float(np.mean([got[t][0]["batch"] for t in tks]))
    for q in range(burst):
        give(bsel[q], bufs[q])
    qps = args.qps if args.qps > 0 else 0.5 * peak

    say(f"serving: burst of {burst} at {peak:.0f} queries/s; Poisson stream of {Q} next")
    # Poisson arrivals
    gsel, srcs = draw(Q)
    gaps = rng.exponential(1.0 / qps, Q)
    check = set(rng.choice(Q, min(Q, args.serve_check), replace=False).tolist())
    for gi in range(G):   # buffers for the expected number in flight, touched now
        need = int(np.ceil(qps * 0.05)) + 64
        while len(pools[gi]) < need:
            pools[gi].append(np.full(graphs[gi].n, -2, dtype=np.int32))
    tickets, held = {}, {}

    def client():
        t = time.perf_counter()
        for q in range(Q):
            t += gaps[q]
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            buf = take(gsel[q])
            with lock:  # the result may come back before the ticket is recorded otherwise
                tk = srv.submit(ids[gsel[q]], srcs[q], out=(buf, None))
                tickets[tk] = q
                held[q] = buf

    results = {}
    t_start = time.perf_counter()
    th = threading.Thread(target=client)
    th.start()
    while len(results) < Q:
        for r, p, _ in srv.poll(wait_ms=50):
            with lock:
                q = tickets[r["ticket"]]
                buf = held.pop(q)
            results[q] = (r, zlib.crc32(p.tobytes()) if q in check else None)
            give(gsel[q], buf)
    th.join()
    t_total = time.perf_counter() - t_start
    load = srv.load()
    srv.close()
    lat = np.array([results[q][0]["latency_ms"] for q in range(Q)])
    qms = np.array([results[q][0]["queue_ms"] for q in range(Q)])
    batch = np.array([results[q][0]["batch"] for q in range(Q)])
    exact = all(results[q][0]["status"] == 0 for q in range(Q))
    # the checked sample against the oracle
    for q in sorted(check):
        g = graphs[gsel[q]]
        ref, _ = oracle.det_bfs(g.offsets, g.adj, srcs[q], threads=host_threads())
        exact &= zlib.crc32(ref.tobytes()) == results[q][1]
    res = {
        "workload": "serve8",
        "graphs": [f"rmat{s} seed {sd} n={g.n} m={g.m}"
                   for (s, sd), g in zip(SERVE_GRAPHS, graphs)],
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p90_ms": round(float(np.percentile(lat, 90)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "throughput_qps": round(Q / t_total, 1),
        "qps_offered": round(qps, 1), "queries": Q,
        "peak_qps_burst": round(peak, 1), "burst_queries": burst,
        "burst_mean_batch": round(burst_batch, 2),
        "mean_batch": round(float(batch.mean()), 2),
        "queue_p50_ms": round(float(np.percentile(qms, 50)), 3),
        "hbm_charged_bytes": load["hbm_in_use"],
        "checked_vs_oracle": len(check), "gen_s": round(gen_s, 1),
        "parity": ("bit-exact vs the oracle on the checked sample" if exact else "MISMATCH"),
    }
    return res, exact


# ---------------------------------------------------------------- multi-graph
MULTI_SMALL = [(sc, sd) for sc in range(10, 18) for sd in (1, 2)]   # 16 graphs, 2^10..2^17


def time_multi(args, graph_list, label):
    """One graph set: one seeded non-isolated source per graph, run as one
    pbbs_bfs_multi call (each job on its own share of one persistent grid;
    jobs on graphs of more than 2^23 edges alone) and as one pbbs_bfs call
    per graph, outputs in HBM while timed; the multi-call outputs are checked
    against the per-graph calls after."""
    import pbbsbench_amd as pb
    rng = np.random.default_rng(args.seed + 7)
    with pb.Context(0) as ctx:
        dgs = [ctx.gen_rmat(sc, 16 * 2 ** sc, 0.57, 0.19, 0.19, seed=sd) for sc, sd in graph_list]
        hosts = [d.download() for d in dgs]
        srcs = [int(rng.choice(np.flatnonzero(h.degrees() > 0))) for h in hosts]
        for d in dgs:
            d.prepare(single=True, batch=False)
        for _ in range(max(args.warmup, 1)):
            ctx.bfs_multi_device(dgs, srcs)
            for d, s in zip(dgs, srcs):
                d.bfs_device(s)
        steps = max(args.steps, 1)
        t0 = time.perf_counter()
        # device time of one call: its shared launch (the same device_ms for every
        # job in it) plus the jobs on graphs of more than 2^23 edges, which
        # pbbs_bfs_multi runs alone, one launch sequence each
        alone = [d.m > 2 ** 23 for d in dgs]

        def call_ms(stats):
            shared = [st["device_ms"] for st, a in zip(stats, alone) if not a]
            return (max(shared) if shared else 0.0) + sum(
                st["device_ms"] for st, a in zip(stats, alone) if a)
        mdev = [call_ms(ctx.bfs_multi_device(dgs, srcs)) for _ in range(steps)]
        multi_ms = (time.perf_counter() - t0) / steps * 1e3
        t0 = time.perf_counter()
        sdev = [sum(d.bfs_device(s)["device_ms"] for d, s in zip(dgs, srcs)) for _ in range(steps)]
        seq_ms = (time.perf_counter() - t0) / steps * 1e3
        out = ctx.bfs_multi(dgs, srcs)
        exact = True
        edges = 0
        for (p, dp, st), d, s in zip(out, dgs, srcs):
            wp, wd, ws = d.bfs(s)
            exact &= bool(np.array_equal(p, wp)) and bool(np.array_equal(dp, wd))
            edges += ws["edges_reached"]
        for d in dgs:
            d.free()
    return {
        "graphs": f"{label}: " + ", ".join(f"rmat{sc}/{sd}" for sc, sd in graph_list),
        "ms_per_launch": round(multi_ms, 4),
        "event_ms_per_launch": round(float(np.mean(mdev)), 4),
        "sequential_wall_ms": round(seq_ms, 4),
        "sequential_event_ms": round(float(np.mean(sdev)), 4),
        "multi_vs_sequential": round(seq_ms / multi_ms, 3),
        "gteps": round(edges / (multi_ms / 1e3) / 1e9, 2),
    }, exact


def run_multi(args):
    """Several graphs of uneven size in one launch, on two sets: 16 small
    RMAT graphs (2^10..2^17 vertices, two seeds each; launch-bound one by
    one) and the eight serving graphs (2^16..2^22)."""
    small, ok1 = time_multi(args, MULTI_SMALL, "small16")
    serve, ok2 = time_multi(args, SERVE_GRAPHS, "serve8")
    exact = ok1 and ok2
    return {"workload": "one query per graph, all graphs in one launch vs one launch "
                        "sequence per graph",
            "small16": small, "serve8": serve,
            "parity": "bit-exact vs one pbbs_bfs per graph" if exact else "MISMATCH"}, exact


# ----------------------------------------------------------------------- ingest
def bench_ingest(args):
    """Graph file -> HBM.  The workload's graph is built on the GPU, downloaded
    and written as an AdjacencyGraph text file and a binary CSR file in a
    temporary directory; a step reads one file into HBM with pbbs_graph_read
    (parse into pinned staging, one upload, the upload's checks).  The files
    sit in the page cache, so this is parse + PCIe + check speed, not disk
    speed.  The host-only reader (pbbs_read_adjacency_graph) is timed beside
    it.  value = text bytes per second into HBM."""
    import ctypes
    import shutil
    import tempfile

    import pbbsbench_amd as pb
    from pbbsbench_amd import _lib
    name = args.workload if args.workload in GPU_BUILDERS else "rmat22"
    wide = WORKLOADS[name][1]
    ctx = pb.Context(0)
    g = GPU_BUILDERS[name](ctx, wide).download()
    need = g.m * 12 + g.n * 20 + (1 << 28)
    tmp = tempfile.mkdtemp(prefix="pbbs_ingest_")
    try:
        if shutil.disk_usage(tmp).free < 2 * need:
            print(f"not enough space in {tmp} for {name}", file=sys.stderr)
            sys.exit(2)
        text, binf = os.path.join(tmp, "g.adj"), os.path.join(tmp, "g.csr")
        t0 = time.perf_counter()
        pb.write_graph_to_file(g, text)
        write_s = time.perf_counter() - t0
        pb.write_graph_binary(g, binf)
        sizes = {"text": os.path.getsize(text), "binary": os.path.getsize(binf)}
        want = (zlib.crc32(g.offsets.tobytes()), zlib.crc32(g.adj.tobytes()))
        res = {}
        for kind, path in (("text", text), ("binary", binf)):
            times = []
            for i in range(max(args.warmup, 1) + max(args.steps, 1)):
                t0 = time.perf_counter()
                dg = ctx.read_graph(path, wide_offsets=wide, assume_symmetric=True)
                dt = time.perf_counter() - t0
                if i == 0:
                    d = dg.download()
                    got = (zlib.crc32(d.offsets.tobytes()), zlib.crc32(d.adj.tobytes()))
                    if got != want:
                        print(json.dumps({"error": f"{kind} read differs from the graph"}))
                        sys.exit(1)
                    del d
                dg.free()
                if i >= max(args.warmup, 1):
                    times.append(dt)
            res[kind] = float(np.median(times))
        # host-only parse (no upload): the C call, no numpy copy
        host = []
        for _ in range(3):
            c = _lib.Csr()
            t0 = time.perf_counter()
            pb._lib.check(_lib.lib().pbbs_read_adjacency_graph(text.encode(), ctypes.byref(c)),
                          "read")
            host.append(time.perf_counter() - t0)
            _lib.lib().pbbs_csr_free(ctypes.byref(c))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    value = sizes["text"] / res["text"] / 1e9
    result = {
        "metric": "AdjacencyGraph text file -> HBM ingest rate",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(res["text"] * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (GPU generator, page-cached files)",
        "config": {"workload": name, "n": g.n, "m": g.m, "text_bytes": sizes["text"],
                   "binary_bytes": sizes["binary"], "offsets": "int64" if wide else "uint32"},
        "text_to_hbm_s": round(res["text"], 3),
        "binary_to_hbm_s": round(res["binary"], 3),
        "binary_GBps": round(sizes["binary"] / res["binary"] / 1e9, 2),
        "host_text_parse_s": round(float(np.median(host)), 3),
        "text_write_s": round(write_s, 2),
        "host_threads": host_threads(),
        "parity": "bit-exact (crc32 of the downloaded CSR)",
    }
    ctx.close()
    print(json.dumps(result))


def mis_traffic():
    """HBM bytes per set of the rounds and the output, from the committed
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic_mis_rmat24jr.json)."""
    path = os.path.join(ROOT, "profiles", "traffic_mis_rmat24jr.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return (t["round_fetch_bytes_per_set"] + t["round_write_bytes_per_set"] +
            t["output_fetch_bytes_per_set"] + t["output_write_bytes_per_set"])


def st_traffic():
    """HBM bytes per forest of the reserve and commit steps, from the committed
    FETCH_SIZE / WRITE_SIZE summary (profiles/traffic_st_rmat24e.json)."""
    path = os.path.join(ROOT, "profiles", "traffic_st_rmat24e.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return (t["reserve_fetch_bytes_per_forest"] + t["reserve_write_bytes_per_forest"] +
            t["commit_fetch_bytes_per_forest"] + t["commit_write_bytes_per_forest"])


MIS_GRAPH = ("rMatGraph_JR_12_16000000 (one of the three graphs the maximal independent set "
             "benchmark times, maximalIndependentSet/bench/testInputs): 2^24 vertices, 12 * 16M "
             "edge draws, a=.55 b=c=.125, relabelled by a seeded permutation, built in HBM")


def run_mis(args):
    """maximalIndependentSet (incrementalMIS/MIS.C:35-66: the lexicographically
    first set) on MIS_GRAPH: device time per set with the flags left in HBM,
    the GPU check of MISCheck's rule, and, with a CPU budget, the oracle's
    serial restatement timed on the same graph and compared bit for bit."""
    import pbbsbench_amd as pb
    ctx = pb.Context(0)
    t0 = time.time()
    dg = ctx.gen_rmat(24, 12 * 16_000_000, 0.55, 0.125, 0.125, seed=1, permute_seed=2)
    gen_s = time.time() - t0
    n, m = dg.n, dg.m
    for _ in range(max(args.warmup, 1)):
        dg.mis_device()
    dev, rounds, t0 = [], [], time.perf_counter()
    for _ in range(args.steps):
        st = dg.mis_device()
        dev.append(st["device_ms"])
        rounds.append(st["rounds"])
    wall_ms = (time.perf_counter() - t0) / args.steps * 1e3
    flags, st_host = dg.mis()
    chk = dg.check_mis(flags)
    ev = float(np.mean(dev))
    So = 8 if dg.info["offsets_i64"] else 4
    B = So * (n + 1) + 4 * m + 5 * n   # offsets, rows, flags in, int32 flags out
    out = {
        "graph": MIS_GRAPH, "n": n, "m": m, "gen_s": round(gen_s, 3),
        "event_ms_per_set": round(ev, 4),
        "ms_per_step": round(wall_ms, 4),
        "rounds_min_max": [int(min(rounds)), int(max(rounds))],
        "members": int(st_host["members"]),
        "d2h_ms": round(st_host["copy_ms"], 3),
        "check": {"valid": chk["valid"], "bad_edges": chk["bad_edges"],
                  "bad_vertices": chk["bad_vertices"]},
        "roofline": {
            "bound": "hbm", "kernel": "mis_init + mis_round per round + mis_output (one set)",
            "achieved": round(B / (ev / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(B / (ev / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes": int(B),
            "algorithmic_note": ("one pass over offsets, rows and flags, plus the int32 output "
                                 "(the set reads row prefixes only, so this normalises)"),
            "traffic": mis_traffic(),
        },
    }
    exact = chk["valid"] == 1
    if args.cpu_seconds > 0:
        import oracle  # the checker and the CPU baseline; never the measured product
        g = dg.download()
        times, ref = [], None
        t1 = time.perf_counter()
        while len(times) < 3 and (not times or time.perf_counter() - t1 + times[-1]
                                  < args.cpu_seconds):
            t2 = time.perf_counter()
            ref = oracle.mis(g.offsets, g.adj)
            times.append(time.perf_counter() - t2)
        same = bool(np.array_equal(ref.astype(np.int32), flags))
        exact &= same
        out["cpu_baseline"] = {
            "value": round(float(np.median(times)) * 1e3, 2), "unit": "ms per set", "cores": 1,
            "kind": "port",
            "sample": (f"the whole set on the same graph, {len(times)} runs of the oracle's "
                       f"serial C restatement of incrementalMIS's rule; CPU: {cpu_info()}"),
        }
        out["parity"] = "bit-exact flags vs oracle" if same else "MISMATCH"
    dg.free()
    ctx.close()
    say(f"mis: {out['event_ms_per_set']} ms per set (device), rounds {out['rounds_min_max']}")
    return out, exact


ST_GRAPH = ("rMatGraph_E_12_16000000's shape (one of the three inputs of the spanning-forest "
            "benchmark, spanningForest/bench/testInputs): 2^24 vertices, 12 * 16M RMAT draws "
            "(a=.55 b=c=.125), relabelled; one (u, v) per undirected edge in a seeded shuffled "
            "order, as the unordered edge files are written; built in HBM")


def run_st(args):
    """spanningForest (incrementalST/ST.C:35-71 under speculative_for's round
    schedule) on ST_GRAPH with the edges and the forest in HBM: device time per
    forest; with a CPU budget, the oracle's serial restatement of the same
    schedule timed on the same edges and compared bit for bit."""
    import pbbsbench_amd as pb
    ctx = pb.Context(0)
    t0 = time.time()
    dg = ctx.gen_rmat(24, 12 * 16_000_000, 0.55, 0.125, 0.125, seed=1, permute_seed=2)
    n = dg.n
    e = dg.edges_device(shuffle_seed=3)
    dg.free()
    gen_s = time.time() - t0
    try:
        for _ in range(max(args.warmup, 1)):
            ctx.spanning_forest_device(n, e.m, e.u, e.v)
        dev, t0 = [], time.perf_counter()
        for _ in range(args.steps):
            cnt, st = ctx.spanning_forest_device(n, e.m, e.u, e.v)
            dev.append(st["device_ms"])
        wall_ms = (time.perf_counter() - t0) / args.steps * 1e3
        # ndST's contract (any spanning forest) on the same edges
        ctx.spanning_forest_device(n, e.m, e.u, e.v, any_forest=True)
        adev = [ctx.spanning_forest_device(n, e.m, e.u, e.v, any_forest=True)[1]["device_ms"]
                for _ in range(args.steps)]
        u, v = e.download()
    finally:
        e.free()
    ev = float(np.mean(dev))
    Bst = 8 * int(u.size) + 12 * n   # each edge read once; parents, R and output once per vertex
    out = {
        "graph": ST_GRAPH, "n": n, "m": int(u.size), "gen_s": round(gen_s, 3),
        "event_ms_per_forest": round(ev, 4),
        "ms_per_step": round(wall_ms, 4),
        "rounds": int(st["rounds"]), "forest_edges": int(cnt),
        "medges_per_s": round(u.size / (ev / 1e3) / 1e6, 1),
        "roofline": {
            "bound": "hbm", "kernel": "st_reserve + st_commit + st_schedule + st_pack per round",
            "achieved": round(Bst / (ev / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(Bst / (ev / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes": Bst,
            "algorithmic_note": ("each edge read once and each vertex's parent, reservation and "
                                 "output once; the schedule's finds re-read parents at random"),
            "traffic": st_traffic(),
        },
        "any_forest": {"contract": "any spanning forest (spanningForest/ndST/ST.C)",
                       "event_ms_per_forest": round(float(np.mean(adev)), 4)},
    }
    exact = True
    if args.cpu_seconds > 0:
        import oracle  # the checker and the CPU baseline; never the measured product
        f, _ = ctx.spanning_forest(n, u, v)
        t1 = time.perf_counter()
        want, rounds = oracle.st(n, u, v)
        cpu_s = time.perf_counter() - t1
        same = bool(np.array_equal(f, want)) and rounds == st["rounds"]
        exact = same and oracle.check_st(n, u, v, f) == 0
        fa, _ = ctx.spanning_forest(n, u, v, any_forest=True)
        any_ok = oracle.check_st(n, u, v, fa) == 0
        out["any_forest"]["check"] = "STCheck rule holds" if any_ok else "INVALID"
        exact = exact and any_ok
        out["cpu_baseline"] = {
            "value": round(cpu_s * 1e3, 1), "unit": "ms per forest", "cores": 1, "kind": "port",
            "sample": ("one forest on the same edges by the oracle's serial C restatement of "
                       f"incrementalST and its round schedule; CPU: {cpu_info()}"),
        }
        out["parity"] = ("bit-exact forest ids and round count vs oracle, STCheck rule holds"
                         if exact else "MISMATCH")
    ctx.close()
    say(f"st: {out['event_ms_per_forest']} ms per forest (device), {out['rounds']} rounds")
    return out, exact


def line(base, **over):
    """A stand-alone JSON line for the batch / serve modes."""
    out = {"n_gpus": 1, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic (seeded generator in libpbbs_bfs)"}
    out.update(base)
    out.update(over)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="rmat24", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=1, help="source-selection seed")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="budget of the CPU-baseline timed rounds after its 1 s warm-up (0 = skip)")
    ap.add_argument("--td-to-bu-div", type=int, default=None)
    ap.add_argument("--bu-to-td-div", type=int, default=None)
    ap.add_argument("--mode", default="single",
                    choices=["single", "batch", "serve", "multi", "ingest", "mis", "st"],
                    help="single: one-source BFS with the batched, serving, multi-graph, "
                         "independent-set and spanning-forest parts nested in its line "
                         "(default); batch: 64 sources "
                         "in one traversal vs 64 sequential runs on RMAT 2^22 alone; serve: "
                         "Poisson query stream over 8 resident RMAT graphs alone; ingest: graph "
                         "file (text / binary) into HBM; mis: maximal independent set on "
                         "rMatGraph_JR_12_16000000 alone; st: spanning forest of the "
                         "rMatGraph_E_12_16000000 shape alone")
    ap.add_argument("--no-batched", action="store_true", help="single: skip the batched part")
    ap.add_argument("--no-serving", action="store_true", help="single: skip the serving part")
    ap.add_argument("--no-any-parent", action="store_true",
                    help="single: skip timing the any-parent (backForwardBFS) contract")
    ap.add_argument("--no-mis", action="store_true",
                    help="single: skip the maximal-independent-set part")
    ap.add_argument("--no-st", action="store_true",
                    help="single: skip the spanning-forest part")
    ap.add_argument("--no-multi", action="store_true",
                    help="single: skip the several-graphs-in-one-launch part")
    ap.add_argument("--no-tuning-table", action="store_true",
                    help="run on the library defaults instead of the committed tuning table")
    ap.add_argument("--qps", type=float, default=0.0,
                    help="serve: Poisson arrival rate (0 = half the measured burst peak)")
    ap.add_argument("--queries", type=int, default=2000, help="serve: queries in the stream")
    ap.add_argument("--serve-check", type=int, default=48,
                    help="serve: queries of the stream compared with the oracle")
    ap.add_argument("--spread", type=int, default=16,
                    help="single: also time this many other seeded non-isolated sources once "
                         "the timed region is over and report their spread (0 = skip)")
    ap.add_argument("--host-gen", action="store_true",
                    help="build the graphs with the host generators (default: on the GPU)")
    ap.add_argument("--tune", action="append", default=[], metavar="FIELD=VALUE",
                    help="single and batch: set a pbbs_tuning field (e.g. persist=0); "
                         "repeatable; replaces the tuning table")
    args = ap.parse_args()
    tune = []
    for kv in args.tune:
        k, _, v = kv.partition("=")
        tune.append((k.strip(), int(v)))
    args.tune = tune
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); this build has no "
              "multi-GPU path: batched and serving work shard by replicating the graph per GPU, "
              "which adds no figure beyond the one-GPU value.", file=sys.stderr)
        sys.exit(2)

    if args.mode == "ingest":
        return bench_ingest(args)
    if args.mode == "batch":
        res, exact = run_batch(args, nested=False)
        print(json.dumps(line({
            "metric": "batched-64-source aggregate GTEPS (vs 64 sequential runs)",
            "value": res["value_gteps"], "unit": "GTEPS", "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["ms_per_batch"],
            "higher_is_better": True, "config": res,
            "roofline": res["roofline"], "parity": res["parity"]},
            **({"cpu_baseline": res.pop("cpu_baseline")} if "cpu_baseline" in res else {}))))
        sys.exit(0 if exact else 1)
    if args.mode == "multi":
        res, exact = run_multi(args)
        print(json.dumps(line({
            "metric": "several graphs of uneven size in one launch vs one launch sequence each",
            "value": res["small16"]["multi_vs_sequential"], "unit": "x", "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["small16"]["ms_per_launch"],
            "higher_is_better": True, "config": res, "parity": res["parity"]})))
        sys.exit(0 if exact else 1)
    if args.mode == "mis":
        res, exact = run_mis(args)
        print(json.dumps(line({
            "metric": "maximal independent set on rMatGraph_JR_12_16000000: ms per set",
            "value": res["event_ms_per_set"], "unit": "ms", "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
            "higher_is_better": False, "config": res,
            "parity": res.get("parity", "GPU check only")},
            **({"cpu_baseline": res.pop("cpu_baseline")} if "cpu_baseline" in res else {}))))
        sys.exit(0 if exact else 1)
    if args.mode == "st":
        res, exact = run_st(args)
        print(json.dumps(line({
            "metric": "spanning forest of the rMatGraph_E_12_16000000 shape: ms per forest",
            "value": res["event_ms_per_forest"], "unit": "ms", "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
            "higher_is_better": False, "config": res,
            "parity": res.get("parity", "not checked (no CPU budget)")},
            **({"cpu_baseline": res.pop("cpu_baseline")} if "cpu_baseline" in res else {}))))
        sys.exit(0 if exact else 1)
    if args.mode == "serve":
        res, exact = run_serve(args, nested=False)
        print(json.dumps(line({
            "metric": "BFS serving: p50/p99 latency (arrival -> parents in host memory)",
            "value": res["p99_ms"], "unit": "ms (p99)", "steps": res["queries"],
            "warmup": len(SERVE_GRAPHS), "ms_per_step": round(1e3 / res["throughput_qps"], 4),
            "higher_is_better": False, "config": res, "parity": res["parity"]})))
        sys.exit(0 if exact else 1)

    result, exact = run_single(args)
    if not args.no_batched:
        result["batched"], ok = run_batch(args, nested=True)
        exact &= ok
    if not args.no_serving:
        result["serving"], ok = run_serve(args, nested=True)
        exact &= ok
    if not args.no_multi:
        say("multi-graph launch")
        result["multi_graph"], ok = run_multi(args)
        exact &= ok
    if not args.no_mis:
        say("maximal independent set")
        result["mis"], ok = run_mis(args)
        exact &= ok
    if not args.no_st:
        say("spanning forest")
        result["spanning_forest"], ok = run_st(args)
        exact &= ok
    print(json.dumps(result))
    if not exact:
        sys.exit(1)


if __name__ == "__main__":
    main()
