"""pbbs_bfs_noalloc, the allocation-free variant of pbbs_bfs: its working
memory is the library's static scratch array, its outputs are the caller's
device buffers.  Every output must equal pbbs_bfs's (and the oracle's) bit for
bit, on two calls in a row with the second on a smaller graph, and the calls
must not take device memory."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import pbbsbench_amd as pb
from pbbsbench_amd import _lib

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        root = os.environ.get("ROCM_PATH", "/opt/rocm")
        _hip = C.CDLL(os.path.join(root, "lib", "libamdhip64.so"))
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _hip


def say(*a):
    print("[case]", *a, flush=True)


class DevBuf:
    """n int32 entries of device memory (the caller's output buffer)."""

    def __init__(self, n):
        self.n = n
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), n * 4) == 0

    def get(self, n):
        out = np.empty(n, dtype=np.int32)
        assert hip().hipMemcpy(out.ctypes.data, self.p, n * 4, 2) == 0   # device -> host
        return out

    def free(self):
        hip().hipFree(self.p)


def free_bytes():
    f, t = C.c_size_t(), C.c_size_t()
    assert hip().hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def noalloc(ctx, dg, src, par, dep):
    st = _lib.BfsStats()
    rc = _lib.lib().pbbs_bfs_noalloc(ctx._h, dg._h, int(src), par.p, dep.p, C.byref(st))
    return rc, st.as_dict()


def case_graphs():
    return [
        ("rmat16", pb.rmat(16, 16 * 2 ** 16, 0.57, 0.19, 0.19, seed=5), False),
        ("torus24", pb.grid3d(24), True),
        ("rmat11", pb.rmat(11, 8 * 2 ** 11, 0.5, 0.1, 0.1, seed=9), True),
    ]


def test_noalloc_equals_pbbs_bfs_two_calls_second_smaller():
    graphs = case_graphs()
    nmax = max(g.n for _, g, _ in graphs)
    with pb.Context(0) as ctx:
        par, dep = DevBuf(nmax), DevBuf(nmax)
        try:
            prev = None
            for name, g, wide in graphs:
                dg = ctx.upload(g, wide_offsets=wide)
                dg.prepare(single=True, batch=False)
                src = int(np.flatnonzero(g.degrees() > 0)[len(name)])
                want_p, want_d, _ = dg.bfs(src)   # the existing entry point
                say(f"{name} n={g.n} src={src} (previous n={prev})")
                rc, st = noalloc(ctx, dg, src, par, dep)
                assert rc == 0, _lib.lib().pbbs_last_error()
                assert np.array_equal(par.get(g.n), want_p)
                assert np.array_equal(dep.get(g.n), want_d)
                op, od = oracle.serial_bfs(g.offsets, g.adj, src)
                assert np.array_equal(want_p, op) and np.array_equal(want_d, od)
                assert st["reached"] == int((want_p != -1).sum())
                if prev is not None:
                    assert g.n < prev   # the second and third calls are smaller
                prev = g.n
                dg.free()
        finally:
            par.free()
            dep.free()


def test_noalloc_takes_no_device_memory():
    g = pb.rmat(20, 16 * 2 ** 20, 0.57, 0.19, 0.19, seed=3)
    src = int(np.flatnonzero(g.degrees() > 0)[11])
    with pb.Context(0) as ctx:
        dg = ctx.upload(g)
        dg.prepare(single=True, batch=False)
        par, dep = DevBuf(g.n), DevBuf(g.n)
        try:
            rc, _ = noalloc(ctx, dg, src, par, dep)   # first call: kernels load
            assert rc == 0, _lib.lib().pbbs_last_error()
            before = free_bytes()
            for _ in range(3):
                rc, _ = noalloc(ctx, dg, src, par, dep)
                assert rc == 0
            after = free_bytes()
            say(f"rmat20 free bytes before {before} after {after}")
            # pbbs_bfs on this fresh context would allocate ~13 n bytes (13.6 MB) here
            assert before - after < (2 << 20)
            want_p, want_d, _ = dg.bfs(src)
            assert np.array_equal(par.get(g.n), want_p)
            assert np.array_equal(dep.get(g.n), want_d)
        finally:
            par.free()
            dep.free()
        dg.free()


def test_noalloc_rejects_what_it_cannot_serve():
    with pb.Context(0) as ctx:
        lib = _lib.lib()
        g = pb.grid3d(8)
        dg = ctx.upload(g)
        par, dep = DevBuf(g.n), DevBuf(g.n)
        try:
            st = _lib.BfsStats()
            # not prepared: the row-top records and share bounds would need allocating
            rc = lib.pbbs_bfs_noalloc(ctx._h, dg._h, 0, par.p, dep.p, C.byref(st))
            assert rc == _lib.PBBS_ERR_STATE and b"prepare" in lib.pbbs_last_error()
            dg.prepare(single=True, batch=False)
            # outputs are required device buffers
            assert lib.pbbs_bfs_noalloc(ctx._h, dg._h, 0, None, dep.p, None) == \
                _lib.PBBS_ERR_INVALID
            assert lib.pbbs_bfs_noalloc(ctx._h, dg._h, g.n, par.p, dep.p, None) == \
                _lib.PBBS_ERR_INVALID
            assert lib.pbbs_bfs_noalloc(ctx._h, dg._h, 0, par.p, dep.p, None) == 0
        finally:
            par.free()
            dep.free()
        dg.free()
        # beyond 2^25 vertices: a clear limit error before anything runs
        big = ctx.gen_grid3d(323)   # 33,698,267 vertices
        bp, bd = DevBuf(big.n), DevBuf(big.n)
        try:
            rc = lib.pbbs_bfs_noalloc(ctx._h, big._h, 0, bp.p, bd.p, None)
            assert rc == _lib.PBBS_ERR_LIMIT and b"33554432" in lib.pbbs_last_error()
        finally:
            bp.free()
            bd.free()
        big.free()
