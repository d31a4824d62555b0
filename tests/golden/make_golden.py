"""Writes the known-answer BFS fixtures in tests/golden/.

The reference holds no golden vectors for this path (its checker only tests
validity), so each fixture here is a small graph whose BFS tree under the
deterministicBFS rule -- parent[v] = the largest neighbour one level above v
(deterministicBFS/BFS.C:65-74, common/atomics.h:75-81), depth of an unreached
vertex = n (bench/BFSCheck.C:55) -- follows from a formula written out per
family below.  Nothing here runs a BFS.  Shapes follow the reference's
generator families where it has them (star, comb, binary tree, n/2 two-vertex
components, 3D torus).

Files: <name>.adj (PBBS AdjacencyGraph), <name>.parent and <name>.depth
(sequenceInt), manifest.json (source vertex and whether the graph is
symmetric).  Run:  python tests/golden/make_golden.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def csr(n, edges, symmetric=True):
    s = set()
    for u, v in edges:
        if symmetric:
            if u == v:
                continue
            s.add((u, v))
            s.add((v, u))
        else:
            s.add((u, v))
    rows = [[] for _ in range(n)]
    for u, v in sorted(s):
        rows[u].append(v)
    off, adj = [0], []
    for r in rows:
        adj.extend(r)
        off.append(len(adj))
    return off, adj


def write_graph(path, off, adj):
    n = len(off) - 1
    with open(path, "w") as f:
        f.write("AdjacencyGraph\n%d\n%d\n" % (n, len(adj)))
        for x in off[:-1]:
            f.write("%d\n" % x)
        for x in adj:
            f.write("%d\n" % x)


def write_seq(path, a):
    with open(path, "w") as f:
        f.write("sequenceInt\n")
        for x in a:
            f.write("%d\n" % x)


def fixtures():
    fx = []

    # path 0-1-...-(n-1) from the middle: parents point toward the source
    n, s = 9, 4
    fx.append(("path9_mid", n, [(i, i + 1) for i in range(n - 1)], s, True,
               [i + 1 if i < s else (i - 1 if i > s else s) for i in range(n)],
               [abs(i - s) for i in range(n)]))

    # complete graph K6 from 2: everyone's parent is 2
    n, s = 6, 2
    fx.append(("complete6", n, [(u, v) for u in range(n) for v in range(u + 1, n)], s, True,
               [s] * n, [0 if i == s else 1 for i in range(n)]))

    # star with centre 0 and 7 leaves, from leaf 5: centre at 1, leaves at 2 via 0
    n, s = 8, 5
    fx.append(("star8_leaf", n, [(0, i) for i in range(1, n)], s, True,
               [s if i in (0, s) else 0 for i in range(n)],
               [0 if i == s else (1 if i == 0 else 2) for i in range(n)]))

    # complete binary tree of 31 nodes from the root: parent (i-1)//2, depth floor(log2(i+1))
    n, s = 31, 0
    fx.append(("bintree31", n, [((i - 1) // 2, i) for i in range(1, n)], s, True,
               [0] + [(i - 1) // 2 for i in range(1, n)],
               [(i + 1).bit_length() - 1 for i in range(n)]))

    # n/2 two-vertex components (2i, 2i+1), from 6: only 7 is reached
    n, s = 12, 6
    fx.append(("pairs12", n, [(2 * i, 2 * i + 1) for i in range(n // 2)], s, True,
               [s if i in (6, 7) else -1 for i in range(n)],
               [0 if i == 6 else (1 if i == 7 else n) for i in range(n)]))

    # diamond 0-1, 0-2, 1-3, 2-3: vertex 3 has parents 1 and 2 one level up -> 2
    fx.append(("diamond", 4, [(0, 1), (0, 2), (1, 3), (2, 3)], 0, True,
               [0, 0, 0, 2], [0, 1, 1, 2]))

    # layered: source 0 -> level 1 = {1..5}; level 2 = {6..9}, vertex 6+j is
    # joined to level-1 vertices 1..(2+j) (so its largest is 2+j);
    # vertex 10 hangs off 9 and off 6 -> largest is 9.
    edges = [(0, i) for i in range(1, 6)]
    for j in range(4):
        edges += [(6 + j, u) for u in range(1, 3 + j)]
    edges += [(10, 9), (10, 6)]
    fx.append(("layered11", 11, edges, 0, True,
               [0, 0, 0, 0, 0, 0, 2, 3, 4, 5, 9],
               [0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3]))

    # comb: spine 0..4, tooth i+5 under spine vertex i; from spine end 4
    n, s = 10, 4
    e = [(i, i + 1) for i in range(4)] + [(i, i + 5) for i in range(5)]
    fx.append(("comb10", n, e, s, True,
               [i + 1 if i < 4 else (4 if i == 4 else i - 5) for i in range(n)],
               [4 - i if i < 5 else (4 - (i - 5)) + 1 for i in range(n)]))

    # isolated source: only itself
    n, s = 5, 3
    fx.append(("isolated_src", n, [(0, 1), (1, 2)], s, True,
               [s if i == s else -1 for i in range(n)],
               [0 if i == s else n for i in range(n)]))

    # directed: 0->1, 0->2, 1->3, 2->3, 3->0, 4->3 (4 unreachable from 0).
    # parent[3] = max(1, 2) = 2; 4 unreached.
    fx.append(("directed6", 6, [(0, 1), (0, 2), (1, 3), (2, 3), (3, 0), (4, 3), (3, 5)], 0,
               False, [0, 0, 0, 2, -1, 3], [0, 1, 1, 2, 6, 3]))

    # 4x4x4 torus from 0 (gridGraph.C:53-72): depth = sum of per-axis torus
    # distances; parent = the largest neighbour id one level up.
    d = 4
    loc = lambda i, j, k: (i % d) * d * d + (j % d) * d + (k % d)
    e = []
    for i in range(d):
        for j in range(d):
            for k in range(d):
                l = loc(i, j, k)
                e += [(l, loc(i + 1, j, k)), (l, loc(i, j + 1, k)), (l, loc(i, j, k + 1))]
    ax = lambda x: min(x, d - x)
    dep, par = [], []
    for i in range(d):
        for j in range(d):
            for k in range(d):
                dv = ax(i) + ax(j) + ax(k)
                dep.append(dv)
                best = -1
                for di, dj, dk in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1),
                                   (0, 0, -1)]:
                    ni, nj, nk = (i + di) % d, (j + dj) % d, (k + dk) % d
                    if ax(ni) + ax(nj) + ax(nk) == dv - 1:
                        best = max(best, loc(ni, nj, nk))
                par.append(best)
    par[0] = 0
    fx.append(("torus4", d ** 3, e, 0, True, par, dep))
    return fx


def main():
    manifest = {}
    for name, n, edges, s, sym, parent, depth in fixtures():
        off, adj = csr(n, edges, symmetric=sym)
        assert len(parent) == n and len(depth) == n
        write_graph(os.path.join(HERE, name + ".adj"), off, adj)
        write_seq(os.path.join(HERE, name + ".parent"), parent)
        write_seq(os.path.join(HERE, name + ".depth"), depth)
        manifest[name] = {"source": s, "symmetric": sym, "n": n, "m": len(adj)}
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", len(manifest), "fixtures")


if __name__ == "__main__":
    main()
