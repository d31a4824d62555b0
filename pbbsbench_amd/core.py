"""Python face of the C ABI, shaped like the reference's benchmark interface.

Reference (benchmarks/breadthFirstSearch/bench/BFS.h:4-11):

    using vertexId = int; using edgeId = uint;
    parlay::sequence<vertexId> BFS(vertexId start, const Graph& G, bool verbose);

and its harness (bench/BFSTime.C:51-61) reads the graph with
readGraphFromFile (common/graphIO.h:191-215) and writes parents with
writeSequenceToFile (common/sequenceIO.h:189-193).  Here:

    G  = read_graph_from_file(path)        # HostGraph (numpy CSR)
    ctx = Context(0)
    dg = ctx.upload(G)                     # DeviceGraph, resident in HBM
    parents = BFS(start, dg)               # numpy int32[n], -1 = unvisited
    write_sequence_to_file(parents, path)

Every call goes through libpbbs_bfs.so; nothing here computes a BFS on the CPU.
"""
import ctypes as C
import os
import threading
import time

import numpy as np

from . import _lib
from ._lib import check, lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class HostGraph:
    """Host CSR: int64 offsets[n+1] (offsets[n] == m), int32 adj[m]."""

    def __init__(self, offsets, adj):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.adj = np.ascontiguousarray(adj, dtype=np.int32)
        if self.offsets.ndim != 1 or self.offsets.size < 2:
            raise ValueError("offsets must hold n+1 >= 2 entries")
        if int(self.offsets[-1]) != self.adj.size:
            raise ValueError("offsets[n] must equal len(adj)")

    @property
    def n(self):
        return self.offsets.size - 1

    @property
    def m(self):
        return self.adj.size

    def numVertices(self):  # common/graph.h:114
        return self.n

    def degrees(self):
        return np.diff(self.offsets)

    def _as_csr(self):
        c = _lib.Csr()
        c.n = self.n
        c.m = self.m
        c.offsets = self.offsets.ctypes.data_as(C.POINTER(C.c_int64))
        c.adj = self.adj.ctypes.data_as(C.POINTER(C.c_int32))
        return c

    @staticmethod
    def _take(csr):
        """Copies a library-allocated pbbs_csr into numpy and frees it."""
        try:
            n, m = csr.n, csr.m
            off = np.ctypeslib.as_array(csr.offsets, shape=(n + 1,)).copy()
            adj = (np.ctypeslib.as_array(csr.adj, shape=(m,)).copy() if m > 0
                   else np.zeros(0, dtype=np.int32))
        finally:
            lib().pbbs_csr_free(C.byref(csr))
        return HostGraph(off, adj)


# ------------------------------------------------------------ constructors
def from_edges(n, u, v, symmetrise=True):
    """CSR from an edge list; self-loops dropped when symmetrising, duplicates
    removed (common/graphUtils.h:54-94)."""
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    if u.shape != v.shape:
        raise ValueError("u and v differ in length")
    c = _lib.Csr()
    check(lib().pbbs_csr_from_edges(n, u.size, _ptr(u), _ptr(v), 1 if symmetrise else 0,
                                    C.byref(c)), "csr_from_edges")
    return HostGraph._take(c)


def rmat(scale, draws, a, b, c, seed=1, permute_seed=0):
    """Symmetric RMAT graph of 2**scale vertices (testData/graphData/rMatGraph.C:33-79)."""
    out = _lib.Csr()
    check(lib().pbbs_gen_rmat(scale, draws, a, b, c, seed, permute_seed, C.byref(out)),
          "gen_rmat")
    return HostGraph._take(out)


def permute(g, seed):
    """g with its vertices relabelled by a seeded random permutation, rows
    sorted (the unordered graphs, e.g. 3Dgrid_JR and randLocalGraph_JR, are
    written this way: common/graphUtils.h:166-189)."""
    src = g._as_csr()
    out = _lib.Csr()
    check(lib().pbbs_csr_permute(C.byref(src), int(seed), C.byref(out)), "csr_permute")
    return HostGraph._take(out)


def grid3d(d):
    """d x d x d torus, degree 6 (testData/graphData/gridGraph.C:53-72)."""
    out = _lib.Csr()
    check(lib().pbbs_gen_grid3d(d, C.byref(out)), "gen_grid3d")
    return HostGraph._take(out)


def randlocal(n, k=10, seed=1):
    """Vertex i draws k neighbours (i +- floor(n**u)) mod n, symmetrised."""
    out = _lib.Csr()
    check(lib().pbbs_gen_randlocal(n, k, seed, C.byref(out)), "gen_randlocal")
    return HostGraph._take(out)


# ------------------------------------------------------------------- text I/O
def read_graph_from_file(path):
    """PBBS AdjacencyGraph reader (common/graphIO.h:191-215)."""
    out = _lib.Csr()
    check(lib().pbbs_read_adjacency_graph(str(path).encode(), C.byref(out)),
          "readGraphFromFile")
    return HostGraph._take(out)


def write_graph_to_file(g, path):
    """PBBS AdjacencyGraph writer (common/graphIO.h:78-105)."""
    c = g._as_csr()
    check(lib().pbbs_write_adjacency_graph(str(path).encode(), C.byref(c)), "writeGraphToFile")


def write_graph_binary(g, path):
    """Writes the binary CSR file ("PBBSCSR1" header, int64 offsets, int32
    adjacency); no reference counterpart (the reference has text formats only)."""
    c = g._as_csr()
    check(lib().pbbs_write_graph_binary(str(path).encode(), C.byref(c)), "write_graph_binary")


def read_graph_binary(path):
    """Reads a file written by write_graph_binary."""
    out = _lib.Csr()
    check(lib().pbbs_read_graph_binary(str(path).encode(), C.byref(out)), "read_graph_binary")
    return HostGraph._take(out)


def write_sequence_to_file(a, path):
    """sequenceInt writer (common/sequenceIO.h:189-193, common/IO.h:151-166)."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    check(lib().pbbs_write_sequence_int(str(path).encode(), _ptr(a), a.size),
          "writeSequenceToFile")


def read_int_seq_from_file(path):
    """sequenceInt reader (common/IO.h:217-229)."""
    p = C.POINTER(C.c_int32)()
    n = C.c_int64()
    check(lib().pbbs_read_sequence_int(str(path).encode(), C.byref(p), C.byref(n)),
          "readIntSeqFromFile")
    try:
        return (np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value
                else np.zeros(0, dtype=np.int32))
    finally:
        lib().pbbs_free(C.cast(p, C.c_void_p))


def read_edge_array_from_file(path):
    """EdgeArray reader (common/graphIO.h:147-166): (n, u, v), n = the largest
    id + 1."""
    pu, pv = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    n, m = C.c_int64(), C.c_int64()
    check(lib().pbbs_read_edge_array(str(path).encode(), C.byref(n), C.byref(m), C.byref(pu),
                                     C.byref(pv)), "readEdgeArrayFromFile")
    try:
        if m.value == 0:
            return n.value, np.zeros(0, np.int32), np.zeros(0, np.int32)
        return (n.value, np.ctypeslib.as_array(pu, shape=(m.value,)).copy(),
                np.ctypeslib.as_array(pv, shape=(m.value,)).copy())
    finally:
        lib().pbbs_free(C.cast(pu, C.c_void_p))
        lib().pbbs_free(C.cast(pv, C.c_void_p))


def write_edge_array_to_file(u, v, path):
    """EdgeArray writer (common/graphIO.h:136-138)."""
    eu = np.ascontiguousarray(u, dtype=np.int32)
    ev = np.ascontiguousarray(v, dtype=np.int32)
    if eu.shape != ev.shape:
        raise ValueError("u and v differ in length")
    check(lib().pbbs_write_edge_array(str(path).encode(), eu.size, _ptr(eu), _ptr(ev)),
          "writeEdgeArrayToFile")


# ------------------------------------------------------------ device objects
class Context:
    """One GPU: a stream, a BFS workspace and captured level graphs."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().pbbs_ctx_create(device, C.byref(h)), "pbbs_ctx_create")
        self._h = h
        self.device = device

    def close(self):
        if self._h:
            lib().pbbs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def tuning(self):
        t = _lib.Tuning()
        check(lib().pbbs_tuning_get(self._h, C.byref(t)), "tuning_get")
        return t

    def set_tuning(self, **kw):
        t = self.tuning
        for k, v in kw.items():
            setattr(t, k, v)
        check(lib().pbbs_tuning_set(self._h, C.byref(t)), "tuning_set")

    def outputs(self, batched=False):
        """Device addresses (ints) of the context's own parent and depth buffers,
        as the last BFS (or batched BFS) with device outputs left them."""
        p, d = C.c_void_p(), C.c_void_p()
        check(lib().pbbs_ctx_outputs(self._h, 1 if batched else 0, C.byref(p), C.byref(d)),
              "ctx_outputs")
        return p.value, d.value

    def upload(self, g, wide_offsets=False, assume_symmetric=False):
        """Copies a HostGraph into HBM.  wide_offsets keeps int64 offsets;
        otherwise they are narrowed to uint32 (the reference's edgeId)."""
        flags = 0
        if wide_offsets:
            off = g.offsets
            flags |= _lib.PBBS_GRAPH_OFFSETS_I64
        else:
            if g.m >= 2**32:
                raise ValueError("m >= 2^32 needs wide offsets")
            off = g.offsets.astype(np.uint32)
        if assume_symmetric:
            flags |= _lib.PBBS_GRAPH_ASSUME_SYMMETRIC
        h = C.c_void_p()
        adj = g.adj if g.m > 0 else np.zeros(1, dtype=np.int32)
        check(lib().pbbs_graph_upload(self._h, g.n, g.m, _ptr(off), _ptr(adj), flags,
                                      C.byref(h)), "pbbs_graph_upload")
        return DeviceGraph(self, h, g.n, g.m)


    def read_graph(self, path, wide_offsets=False, assume_symmetric=False):
        """Reads an AdjacencyGraph text file or a binary CSR file (told apart
        by its magic) straight into HBM through pinned staging: the
        readGraphFromFile + upload pair (common/graphIO.h:191-215)."""
        flags = _lib.PBBS_GRAPH_OFFSETS_I64 if wide_offsets else 0
        if assume_symmetric:
            flags |= _lib.PBBS_GRAPH_ASSUME_SYMMETRIC
        h = C.c_void_p()
        check(lib().pbbs_graph_read(self._h, str(path).encode(), flags, C.byref(h)),
              "graph_read")
        g = DeviceGraph(self, h, 0, 0)
        i = g.info
        g.n, g.m = i["n"], i["m"]
        return g

    def bfs_multi(self, graphs, sources, depth=True):
        """Single-source BFS on each (DeviceGraph, source) pair in one
        persistent launch (pbbs_bfs_multi).  Returns a list of (parent, depth
        or None, stats) in job order; each equals graphs[j].bfs(sources[j])."""
        J = len(graphs)
        if J != len(sources):
            raise ValueError("one source per graph")
        par = [np.empty(g.n, dtype=np.int32) for g in graphs]
        dep = [np.empty(g.n, dtype=np.int32) for g in graphs] if depth else None
        gh = (C.c_void_p * J)(*[g._h for g in graphs])
        src = np.ascontiguousarray(sources, dtype=np.int32)
        pp = (C.c_void_p * J)(*[a.ctypes.data for a in par])
        dp = (C.c_void_p * J)(*[a.ctypes.data for a in dep]) if depth else None
        st = (_lib.BfsStats * J)()
        check(lib().pbbs_bfs_multi(self._h, J, gh, _ptr(src), pp, dp, 0, st), "BFS multi")
        return [(par[j], dep[j] if depth else None, st[j].as_dict()) for j in range(J)]

    def bfs_multi_device(self, graphs, sources):
        """bfs_multi with the outputs left in HBM (the context's multi-graph
        workspace); returns the per-job stats dicts (device_ms is the whole
        launch's)."""
        J = len(graphs)
        if J != len(sources):
            raise ValueError("one source per graph")
        gh = (C.c_void_p * J)(*[g._h for g in graphs])
        src = np.ascontiguousarray(sources, dtype=np.int32)
        st = (_lib.BfsStats * J)()
        check(lib().pbbs_bfs_multi(self._h, J, gh, _ptr(src), None, None, _lib.PBBS_OUT_DEVICE,
                                   st), "BFS multi")
        return [st[j].as_dict() for j in range(J)]

    def spanning_forest(self, n, u, v, any_forest=False):
        """Spanning forest of the edge array (u[i], v[i]) on n vertices: the edge
        ids incrementalST returns (spanningForest/incrementalST/ST.C:35-71), in
        its order; any_forest: some spanning forest (ndST's contract,
        spanningForest/ndST/ST.C).  Returns (ids int32[k], stats dict)."""
        eu = np.ascontiguousarray(u, dtype=np.int32)
        ev = np.ascontiguousarray(v, dtype=np.int32)
        if eu.shape != ev.shape:
            raise ValueError("u and v differ in length")
        out = np.empty(max(int(n) - 1, 1), dtype=np.int32)
        cnt = C.c_int64()
        st = _lib.StStats()
        flags = _lib.PBBS_ST_ANY_FOREST if any_forest else 0
        check(lib().pbbs_spanning_forest(self._h, int(n), eu.size, _ptr(eu), _ptr(ev), flags,
                                         _ptr(out), C.byref(cnt), C.byref(st)), "spanning_forest")
        return out[:cnt.value].copy(), st.as_dict()

    def spanning_forest_device(self, n, m, u_ptr, v_ptr, forest_ptr=None, any_forest=False):
        """spanning_forest() on edge arrays already in HBM (device pointers to m
        int32 each); the ids stay in HBM (forest_ptr, room for n - 1, or the
        context's own buffer when None).  Returns (count, stats dict)."""
        cnt = C.c_int64()
        st = _lib.StStats()
        flags = (_lib.PBBS_EDGES_DEVICE | (_lib.PBBS_OUT_DEVICE if forest_ptr else 0) |
                 (_lib.PBBS_ST_ANY_FOREST if any_forest else 0))
        check(lib().pbbs_spanning_forest(self._h, int(n), int(m), u_ptr, v_ptr, flags,
                                         forest_ptr, C.byref(cnt), C.byref(st)),
              "spanning_forest")
        return cnt.value, st.as_dict()

    def gen_rmat(self, scale, draws, a, b, c, seed=1, permute_seed=0, wide_offsets=False):
        """rmat(...) built on the GPU, straight into HBM; the same graph bit for
        bit as ctx.upload(rmat(...)) (testData/graphData/rMatGraph.C:33-79)."""
        h = C.c_void_p()
        check(lib().pbbs_graph_gen_rmat(self._h, scale, draws, a, b, c, seed, permute_seed,
                                        _lib.PBBS_GRAPH_OFFSETS_I64 if wide_offsets else 0,
                                        C.byref(h)), "graph_gen_rmat")
        g = DeviceGraph(self, h, 0, 0)
        i = g.info
        g.n, g.m = i["n"], i["m"]
        return g

    def gen_grid3d(self, d, wide_offsets=False):
        """grid3d(d) built on the GPU (testData/graphData/gridGraph.C:53-72)."""
        h = C.c_void_p()
        check(lib().pbbs_graph_gen_grid3d(self._h, d,
                                          _lib.PBBS_GRAPH_OFFSETS_I64 if wide_offsets else 0,
                                          C.byref(h)), "graph_gen_grid3d")
        g = DeviceGraph(self, h, 0, 0)
        i = g.info
        g.n, g.m = i["n"], i["m"]
        return g


class DeviceGraph:
    """A CSR resident in HBM."""

    def __init__(self, ctx, handle, n, m):
        self.ctx = ctx
        self._h = handle
        self.n = n
        self.m = m

    def free(self):
        if self._h:
            lib().pbbs_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def numVertices(self):
        return self.n

    def check(self, source, parent, depth=None):
        """GPU checker (bench/BFSCheck.C:59-90 + the max-parent rule): returns
        {valid, cycle, first_bad, bad, max_rule_bad, depth_bad, reached, max_level}."""
        parent = np.ascontiguousarray(parent, dtype=np.int32)
        if parent.size != self.n:
            raise ValueError("parent must hold n entries")
        dep = None
        if depth is not None:
            dep = np.ascontiguousarray(depth, dtype=np.int32)
            if dep.size != self.n:
                raise ValueError("depth must hold n entries")
        r = _lib.CheckResult()
        check(lib().pbbs_check_bfs(self.ctx._h, self._h, int(source), _ptr(parent),
                                   _ptr(dep) if dep is not None else None, 0, C.byref(r)),
              "check_bfs")
        return r.as_dict()

    def check_device(self, source, parent_ptr, depth_ptr=None):
        """check() on outputs already in HBM (device pointers)."""
        r = _lib.CheckResult()
        check(lib().pbbs_check_bfs(self.ctx._h, self._h, int(source), parent_ptr, depth_ptr,
                                   _lib.PBBS_CHECK_DEVICE_INPUT, C.byref(r)), "check_bfs")
        return r.as_dict()

    def edges_device(self, shuffle_seed=0):
        """The undirected edges (u < v) of this symmetric graph as an edge array
        in HBM, in row order or shuffled by a seeded permutation.  Returns a
        DeviceEdges (free it when done)."""
        m = C.c_int64()
        check(lib().pbbs_graph_edges(self.ctx._h, self._h, 0, None, None, 0, C.byref(m)),
              "graph_edges")
        e = DeviceEdges(self.n, m.value)
        try:
            check(lib().pbbs_graph_edges(self.ctx._h, self._h, int(shuffle_seed), e.u, e.v,
                                         e.m, C.byref(m)), "graph_edges")
        except Exception:
            e.free()
            raise
        return e

    def download(self):
        """The resident CSR as a HostGraph (int64 offsets)."""
        c = _lib.Csr()
        check(lib().pbbs_graph_download(self._h, C.byref(c)), "graph_download")
        return HostGraph._take(c)

    def prepare(self, single=True, batch=True):
        """Builds the graph's search data now (pbbs_graph_prepare) rather than
        on its first BFS; returns the call's wall time in ms."""
        what = (_lib.PBBS_PREPARE_SINGLE if single else 0) | (_lib.PBBS_PREPARE_BATCH if batch else 0)
        ms = C.c_double()
        check(lib().pbbs_graph_prepare(self.ctx._h, self._h, what, C.byref(ms)), "graph_prepare")
        return ms.value

    @property
    def info(self):
        i = _lib.GraphInfo()
        check(lib().pbbs_graph_info_get(self._h, C.byref(i)), "graph_info")
        return {name: getattr(i, name) for name, _ in i._fields_ if name != "pad"}

    def bfs(self, source, depth=True, any_parent=False):
        """Returns (parent, depth or None, stats dict); outputs in host memory.
        any_parent: the backForwardBFS contract (backForwardBFS/BFS.C:38-61),
        any neighbour one level up as the parent; depths are the same."""
        parent = np.empty(self.n, dtype=np.int32)
        dep = np.empty(self.n, dtype=np.int32) if depth else None
        st = _lib.BfsStats()
        flags = _lib.PBBS_BFS_ANY_PARENT if any_parent else 0
        check(lib().pbbs_bfs(self.ctx._h, self._h, int(source), _ptr(parent),
                             _ptr(dep) if depth else None, flags, C.byref(st)), "BFS")
        return parent, dep, st.as_dict()

    def bfs_batch(self, sources, depth=True):
        """BFS from each source (at most 64) in one traversal.  Returns
        (parents[k][n], depths[k][n] or None, stats); row i equals bfs(sources[i])."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        k = src.size
        parents = np.empty((k, self.n), dtype=np.int32)
        deps = np.empty((k, self.n), dtype=np.int32) if depth else None
        st = _lib.BfsStats()
        check(lib().pbbs_bfs_batch(self.ctx._h, self._h, _ptr(src), k, _ptr(parents),
                                   _ptr(deps) if depth else None, 0, C.byref(st)), "BFS batch")
        return parents, deps, st.as_dict()

    def bfs_batch_device(self, sources, parent_ptr=None, depth_ptr=None):
        """Batched BFS with outputs left in HBM; returns the stats dict."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        st = _lib.BfsStats()
        check(lib().pbbs_bfs_batch(self.ctx._h, self._h, _ptr(src), src.size, parent_ptr,
                                   depth_ptr, _lib.PBBS_OUT_DEVICE, C.byref(st)), "BFS batch")
        return st.as_dict()

    def profile(self, source, max_levels=4096):
        """Per-level kernel times of one BFS (levels launched one by one).
        Returns (init_ms, [ {level, mode, frontier, frontier_edges, ms}, ... ])."""
        recs = (_lib.LevelProfile * max_levels)()
        cnt = C.c_int32()
        init = C.c_double()
        check(lib().pbbs_bfs_profile(self.ctx._h, self._h, int(source), recs, max_levels,
                                     C.byref(cnt), C.byref(init)), "bfs_profile")
        out = []
        for r in recs[:min(cnt.value, max_levels)]:
            out.append({"level": r.level, "mode": "BU" if r.mode else "TD",
                        "frontier": r.frontier, "frontier_edges": r.frontier_edges, "ms": r.ms})
        return init.value, out

    def mis(self):
        """Maximal independent set (incrementalMIS/MIS.C:35-66: v is in the set
        unless a lower-numbered neighbour is).  Returns (flags int32[n], 1 in
        the set and 2 not, as MISTime.C writes them; stats dict)."""
        flags = np.empty(self.n, dtype=np.int32)
        st = _lib.MisStats()
        check(lib().pbbs_mis(self.ctx._h, self._h, _ptr(flags), 0, C.byref(st)), "MIS")
        return flags, st.as_dict()

    def check_mis(self, flags):
        """MISCheck's rule (maximalIndependentSet/bench/MISCheck.C:37-63) on the
        GPU for host flags (int32[n]); returns the result dict."""
        f = np.ascontiguousarray(flags, dtype=np.int32)
        if f.size != self.n:
            raise ValueError("flags must have n entries")
        r = _lib.MisCheckResult()
        check(lib().pbbs_check_mis(self.ctx._h, self._h, _ptr(f), 0, C.byref(r)), "check_mis")
        return r.as_dict()

    def check_mis_device(self, flags_ptr):
        """check_mis() for flags already in HBM (a device pointer to n int32)."""
        r = _lib.MisCheckResult()
        check(lib().pbbs_check_mis(self.ctx._h, self._h, flags_ptr, _lib.PBBS_CHECK_DEVICE_INPUT,
                                   C.byref(r)), "check_mis")
        return r.as_dict()

    def mis_device(self, flags_ptr=None):
        """mis() with the flags left in HBM (a device pointer to n int32, or
        the context's own buffer when None).  Returns the stats dict."""
        st = _lib.MisStats()
        check(lib().pbbs_mis(self.ctx._h, self._h, flags_ptr, _lib.PBBS_OUT_DEVICE,
                             C.byref(st)), "MIS")
        return st.as_dict()

    def bfs_device(self, source, parent_ptr=None, depth_ptr=None, any_parent=False):
        """BFS with outputs left in HBM (device pointers, or the context's own
        buffers when None).  Returns the stats dict."""
        st = _lib.BfsStats()
        flags = _lib.PBBS_OUT_DEVICE | (_lib.PBBS_BFS_ANY_PARENT if any_parent else 0)
        check(lib().pbbs_bfs(self.ctx._h, self._h, int(source), parent_ptr, depth_ptr,
                             flags, C.byref(st)), "BFS")
        return st.as_dict()


_hip = None


def hip_runtime():
    """The HIP runtime library, for device buffers the caller owns."""
    global _hip
    if _hip is None:
        root = os.environ.get("ROCM_PATH", "/opt/rocm")
        h = C.CDLL(os.path.join(root, "lib", "libamdhip64.so"))
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip = h
    return _hip


class DeviceEdges:
    """An edge array of m (u, v) int32 pairs in HBM on n vertices (the
    reference's edgeArray, resident).  u and v are device addresses."""

    def __init__(self, n, m):
        self.n, self.m = int(n), int(m)
        self._u, self._v = C.c_void_p(), C.c_void_p()
        h = hip_runtime()
        if h.hipMalloc(C.byref(self._u), max(self.m, 1) * 4) != 0 or \
                h.hipMalloc(C.byref(self._v), max(self.m, 1) * 4) != 0:
            self.free()
            raise MemoryError(f"device edge array of {m} entries")

    @property
    def u(self):
        return self._u.value

    @property
    def v(self):
        return self._v.value

    def download(self):
        """(u, v) as host int32 arrays."""
        u = np.empty(self.m, dtype=np.int32)
        v = np.empty(self.m, dtype=np.int32)
        h = hip_runtime()
        if self.m:
            assert h.hipMemcpy(u.ctypes.data, self._u, self.m * 4, 2) == 0
            assert h.hipMemcpy(v.ctypes.data, self._v, self.m * 4, 2) == 0
        return u, v

    def free(self):
        h = hip_runtime()
        for p in (self._u, self._v):
            if p.value:
                h.hipFree(p)
                p.value = None


class Server:
    """Serving loop over resident graphs (pbbs_server_* in include/pbbs_bfs.h).

    submit() queues a BFS and returns a ticket; the output arrays are owned by
    the Server until the ticket's result comes back from poll().  Queries for
    one symmetric graph that wait together are served by one batched
    traversal; each result is identical to a lone Context BFS."""

    def __init__(self, device=0, slots=0, max_batch=0, hbm_budget=0):
        cfg = _lib.ServerConfig()
        cfg.slots = slots
        cfg.max_batch = max_batch
        cfg.hbm_budget = hbm_budget
        h = C.c_void_p()
        check(lib().pbbs_server_create(device, C.byref(cfg), C.byref(h)), "server_create")
        self._h = h
        self._graph_n = {}
        self._outputs = {}
        self._lock = threading.Lock()   # a result can be polled before submit() returns

    def close(self):
        if self._h:
            lib().pbbs_server_destroy(self._h)
            self._h = None
            self._outputs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_graph(self, g, wide_offsets=False, assume_symmetric=False):
        flags = 0
        if wide_offsets:
            off = g.offsets
            flags |= _lib.PBBS_GRAPH_OFFSETS_I64
        else:
            if g.m >= 2**32:
                raise ValueError("m >= 2^32 needs wide offsets")
            off = g.offsets.astype(np.uint32)
        if assume_symmetric:
            flags |= _lib.PBBS_GRAPH_ASSUME_SYMMETRIC
        gid = C.c_int32()
        adj = g.adj if g.m > 0 else np.zeros(1, dtype=np.int32)
        check(lib().pbbs_server_add_graph(self._h, g.n, g.m, _ptr(off), _ptr(adj), flags,
                                          C.byref(gid)), "server_add_graph")
        self._graph_n[gid.value] = g.n
        return gid.value

    def submit(self, graph_id, source, depth=True, out=None):
        """Queues BFS(source) on a graph; returns the ticket.  `out` may give
        (parent, depth-or-None) int32 arrays of the graph's n entries to fill,
        e.g. reused buffers; otherwise fresh arrays are allocated."""
        n = max(self._graph_n.get(graph_id, 1), 1)   # an unknown id is rejected below
        if out is not None:
            parent, dep = out
            for a in (parent, dep):
                if a is not None and (a.dtype != np.int32 or a.size != n
                                      or not a.flags.c_contiguous):
                    raise ValueError(f"output buffers must be contiguous int32[{n}]")
            depth = dep is not None
        else:
            parent = np.empty(n, dtype=np.int32)
            dep = np.empty(n, dtype=np.int32) if depth else None
        t = C.c_uint64()
        with self._lock:
            check(lib().pbbs_server_submit(self._h, graph_id, int(source), _ptr(parent),
                                           _ptr(dep) if depth else None, C.byref(t)),
                  "server_submit")
            self._outputs[t.value] = (parent, dep)
        return t.value

    def poll(self, max_results=256, wait_ms=0):
        """Finished queries: a list of (result dict, parent, depth or None)."""
        buf = (_lib.ServerResult * max_results)()
        cnt = C.c_int32()
        check(lib().pbbs_server_poll(self._h, buf, max_results, C.byref(cnt), wait_ms),
              "server_poll")
        out = []
        with self._lock:
            for r in buf[:cnt.value]:
                parent, dep = self._outputs.pop(r.ticket)
                out.append((r.as_dict(), parent, dep))
        return out

    def drain(self, tickets, timeout_ms=60000):
        """Polls until every ticket in `tickets` has finished; returns
        {ticket: (result, parent, depth)}."""
        want = set(tickets)
        got = {}
        end = time.monotonic() + timeout_ms / 1000.0
        while want:
            if time.monotonic() > end:
                raise TimeoutError(f"{len(want)} queries unfinished")
            for r, p, d in self.poll(wait_ms=100):
                got[r["ticket"]] = (r, p, d)
                want.discard(r["ticket"])
        return got

    def load(self):
        q, r, h = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().pbbs_server_load(self._h, C.byref(q), C.byref(r), C.byref(h)),
              "server_load")
        return {"queued": q.value, "running": r.value, "hbm_in_use": h.value}


def BFS(start, G, verbose=False):
    """Parent sequence of a BFS from `start`: -1 = unvisited, the start
    points to itself (bench/BFS.h:9-11)."""
    parent, _, _ = G.bfs(start, depth=False)
    return parent
