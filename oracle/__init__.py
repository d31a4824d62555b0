"""ORACLE -- test infrastructure only (the checker, never the product).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package.  It loads oracle/liboracle.so, a C restatement of PBBS
deterministicBFS and its checker (see bfs_oracle.c for the reference
citations), and adds numpy closed forms for graph families whose BFS tree is
known in closed form.

Parity pins (what checks each function, so that a plausible mistake fails):
  * det_bfs       == serial_bfs (independent brute force of the max-parent
                     closed form) on random graphs of many seeds and shapes;
                  == the committed golden fixtures (hand-derived trees);
                  == torus_closed_form on d^3 tori.
  * serial_bfs    == golden fixtures; depths == level_number(parents).
  * level_number  == serial_bfs depths; == fixture depths.
  * check_bfs     accepts every oracle output; rejects planted mistakes
                  (wrong level, wrong root, unreached marked reached, cycles).
  * max_rule      zero on oracle outputs; non-zero on planted min-parents.
  * mis           == the lexicographically first maximal independent set found
                     by brute force over all vertex subsets of small random
                     graphs; == closed forms (paths, stars, complete graphs,
                     even tori: the vertices of even coordinate sum).
  * check_mis     accepts every mis output; rejects planted mistakes (two
                  adjacent members, an outsider with no member neighbour).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def _l():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing; run make -C oracle")
        h = C.CDLL(LIB_PATH)
        P = C.c_void_p
        h.oracle_det_bfs.restype = C.c_int64
        h.oracle_det_bfs.argtypes = [C.c_int64, P, P, C.c_int32, P, C.c_int]
        h.oracle_level_number.restype = C.c_int
        h.oracle_level_number.argtypes = [C.c_int64, P, P]
        h.oracle_check_bfs.restype = C.c_int
        h.oracle_check_bfs.argtypes = [C.c_int64, P, P, C.c_int32, P, C.c_int64,
                                       C.POINTER(C.c_int64)]
        h.oracle_serial_bfs.restype = C.c_int
        h.oracle_serial_bfs.argtypes = [C.c_int64, P, P, C.c_int32, P, P]
        h.oracle_max_rule.restype = C.c_int64
        h.oracle_max_rule.argtypes = [C.c_int64, P, P, C.c_int32, P, P]
        h.oracle_mis.argtypes = [C.c_int64, P, P, P]
        h.oracle_check_mis.argtypes = [C.c_int64, P, P, P]
        h.oracle_check_mis.restype = C.c_int64
        h.oracle_st.restype = C.c_int64
        h.oracle_st.argtypes = [C.c_int64, C.c_int64, P, P, C.c_int64, P, P]
        h.oracle_serial_st.restype = C.c_int64
        h.oracle_serial_st.argtypes = [C.c_int64, C.c_int64, P, P, P]
        h.oracle_check_st.restype = C.c_int
        h.oracle_check_st.argtypes = [C.c_int64, C.c_int64, P, P, P, C.c_int64]
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _csr(offsets, adj):
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    a = np.ascontiguousarray(adj, dtype=np.int32)
    if a.size == 0:
        a = np.zeros(1, dtype=np.int32)
    return off, a


def det_bfs(offsets, adj, source, threads=0):
    """deterministicBFS restated (BFS.C:39-78).  Returns (parent, rounds)."""
    off, a = _csr(offsets, adj)
    n = off.size - 1
    parent = np.empty(n, dtype=np.int32)
    rounds = _l().oracle_det_bfs(n, _p(off), _p(a), int(source), _p(parent), int(threads))
    if rounds < 0:
        raise MemoryError("oracle_det_bfs")
    return parent, int(rounds)


def level_number(parent):
    """bench/BFSCheck.C:39-56.  Unreached -> n.  Raises on a parent cycle."""
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    level = np.empty_like(parent)
    rc = _l().oracle_level_number(parent.size, _p(parent), _p(level))
    if rc != 0:
        raise ValueError("level_number: no fixed point (cycle in parents)" if rc > 0
                         else "level_number: out of memory")
    return level


def check_bfs(offsets, adj, source, parents):
    """bench/BFSCheck.C:59-90 with a source parameter.  Returns -1 if valid,
    else the first failing vertex (or -2 for a size mismatch / cycle)."""
    off, a = _csr(offsets, adj)
    parents = np.ascontiguousarray(parents, dtype=np.int32)
    fail = C.c_int64(-1)
    rc = _l().oracle_check_bfs(off.size - 1, _p(off), _p(a), int(source), _p(parents),
                               parents.size, C.byref(fail))
    if rc < 0:
        raise MemoryError("oracle_check_bfs")
    if rc == 0:
        return -1
    return int(fail.value) if fail.value >= 0 else -2


def serial_bfs(offsets, adj, source):
    """Brute force of the closed form: FIFO depths, parent = largest in-neighbour
    one level up.  Returns (parent, depth)."""
    off, a = _csr(offsets, adj)
    n = off.size - 1
    parent = np.empty(n, dtype=np.int32)
    depth = np.empty(n, dtype=np.int32)
    if _l().oracle_serial_bfs(n, _p(off), _p(a), int(source), _p(parent), _p(depth)) != 0:
        raise MemoryError("oracle_serial_bfs")
    return parent, depth


def max_rule_violations(offsets, adj, source, parent, depth):
    """Symmetric graphs: vertices whose parent is not the largest neighbour one
    level up (or whose unreached / root entry is wrong)."""
    off, a = _csr(offsets, adj)
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    depth = np.ascontiguousarray(depth, dtype=np.int32)
    return int(_l().oracle_max_rule(off.size - 1, _p(off), _p(a), int(source), _p(parent),
                                    _p(depth)))


def torus_closed_form(d):
    """BFS from vertex 0 on the d^3 torus of gridGraph.C:53-72, in closed form:
    depth = sum of per-axis torus distances; parent = the largest neighbour id
    whose depth is one less.  No traversal involved."""
    idx = np.arange(d, dtype=np.int64)
    ax = np.minimum(idx, d - idx)
    i, j, k = np.meshgrid(idx, idx, idx, indexing="ij")
    depth = (ax[i] + ax[j] + ax[k]).astype(np.int32)
    vid = (i * d * d + j * d + k)
    parent = np.full(d ** 3, -1, dtype=np.int64).reshape(d, d, d)
    for axis, (di, dj, dk) in enumerate([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0),
                                         (0, 0, 1), (0, 0, -1)]):
        ni, nj, nk = (i + di) % d, (j + dj) % d, (k + dk) % d
        nid = ni * d * d + nj * d + nk
        ok = (ax[ni] + ax[nj] + ax[nk]) == depth - 1
        parent = np.where(ok & (nid > parent), nid, parent)
    parent = parent.reshape(-1).astype(np.int32)
    parent[0] = 0
    del vid
    return parent, depth.reshape(-1)


def mis(offsets, adj):
    """maximalIndependentSet/incrementalMIS/MIS.C:35-66: flags[v] = 1 in the
    lexicographically first maximal independent set, 2 otherwise."""
    off, a = _csr(offsets, adj)
    flags = np.empty(off.size - 1, dtype=np.int32)
    _l().oracle_mis(off.size - 1, _p(off), _p(a), _p(flags))
    return flags


def check_mis(offsets, adj, flags):
    """maximalIndependentSet/bench/MISCheck.C:37-60: -1 if flags mark a maximal
    independent set (1 = member), else the first bad vertex."""
    off, a = _csr(offsets, adj)
    flags = np.ascontiguousarray(flags, dtype=np.int32)
    return int(_l().oracle_check_mis(off.size - 1, _p(off), _p(a), _p(flags)))


def torus_mis_closed_form(d):
    """On the d^3 torus of gridGraph.C:53-72 with d even, every edge joins
    coordinate sums of opposite parity, and each odd vertex has a lower even
    neighbour (one coordinate minus one): the first maximal independent set is
    the even vertices.  Returns its flags (1 / 2)."""
    assert d % 2 == 0
    idx = np.arange(d)
    i, j, k = np.meshgrid(idx, idx, idx, indexing="ij")
    return np.where((i + j + k) % 2 == 0, 1, 2).astype(np.int32).reshape(-1)


def _edges(u, v):
    eu = np.ascontiguousarray(u, dtype=np.int32)
    ev = np.ascontiguousarray(v, dtype=np.int32)
    assert eu.shape == ev.shape
    return eu, ev


def st(n, u, v, granularity=100):
    """spanningForest/incrementalST/ST.C:35-71 under speculative_for's round
    schedule (common/speculative_for.h:30-100).  Returns (forest edge ids in
    the reference's output order, rounds), or (None, rounds) when the
    schedule gives up as the reference does (m < 3 * granularity)."""
    eu, ev = _edges(u, v)
    out = np.empty(max(int(n), 1), dtype=np.int32)
    rounds = C.c_int64()
    k = _l().oracle_st(int(n), eu.size, _p(eu), _p(ev), int(granularity), _p(out),
                       C.byref(rounds))
    return (None if k < 0 else out[:k].copy()), rounds.value


def serial_st(n, u, v):
    """spanningForest/bench/serialST.C: the edges, in order, that join two trees."""
    eu, ev = _edges(u, v)
    out = np.empty(max(eu.size, 1), dtype=np.int32)
    k = _l().oracle_serial_st(int(n), eu.size, _p(eu), _p(ev), _p(out))
    return out[:k].copy()


def check_st(n, u, v, forest):
    """spanningForest/bench/STCheck.C:36-61: 0 if `forest` is a spanning forest
    of the edges (u, v); 1 wrong size, 2 a cycle, 3 an edge id out of range."""
    eu, ev = _edges(u, v)
    f = np.ascontiguousarray(forest, dtype=np.int32)
    return int(_l().oracle_check_st(int(n), eu.size, _p(eu), _p(ev), _p(f), f.size))

