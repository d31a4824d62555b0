"""The C ABI library loads and exports every function include/*.h declares.
No compute call is made here (this suite runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT, has_gpu

HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(pbbs_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("pbbs_ctx_create", "pbbs_graph_upload", "pbbs_bfs", "pbbs_read_adjacency_graph"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from pbbsbench_amd import _lib
    lib = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound, set(declared_functions()) - bound


def test_version_and_error_strings():
    from pbbsbench_amd import _lib
    lib = _lib.lib()
    assert lib.pbbs_version().startswith(b"pbbs-bfs-gfx950")
    assert isinstance(lib.pbbs_last_error(), bytes)


@pytest.mark.skipif(has_gpu(), reason="checks the no-GPU error path")
def test_context_without_gpu_fails_cleanly():
    import pbbsbench_amd as pb
    with pytest.raises(pb.PbbsError):
        pb.Context(0)


def test_null_arguments_rejected():
    from pbbsbench_amd import _lib
    lib = _lib.lib()
    assert lib.pbbs_ctx_create(0, None) == _lib.PBBS_ERR_INVALID
    assert lib.pbbs_graph_upload(None, 1, 0, None, None, 0, None) == _lib.PBBS_ERR_INVALID
    assert lib.pbbs_bfs(None, None, 0, None, None, 0, None) == _lib.PBBS_ERR_INVALID
    assert b"NULL" in lib.pbbs_last_error()


def test_driver_binary_usage():
    exe = os.path.join(ROOT, "pbbsbench_amd", "bin", "BFS")
    assert os.path.exists(exe)
    out = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=30)
    assert "usage:" in out.stdout and "-src" in out.stdout and "-any" in out.stdout


def test_driver_reads_graph_and_rejects_bad_source():
    """The driver parses the AdjacencyGraph file (common/graphIO.h:191-215) and
    rejects a source outside [0, n) before it touches a device; a missing file
    is an error too.  Neither run reaches the GPU."""
    exe = os.path.join(ROOT, "pbbsbench_amd", "bin", "BFS")
    g = os.path.join(ROOT, "tests", "golden", "bintree31.adj")
    out = subprocess.run([exe, "-src", "31", g], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "source 31 outside [0, 31)" in out.stderr
    out = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "no_such.adj")],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "readGraphFromFile" in out.stderr


@pytest.mark.parametrize("name", ["MIS", "ST"])
def test_set_and_forest_drivers_usage_and_bad_files(tmp_path, name):
    """bin/MIS and bin/ST take the reference drivers' command line
    (MISTime.C:48-55, STTime.C:45-52) and fail on a missing or malformed
    input before they touch a device."""
    exe = os.path.join(ROOT, "pbbsbench_amd", "bin", name)
    assert os.path.exists(exe)
    out = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=30)
    assert out.stdout.startswith("usage:") and "[-o <outFile>] [-r <rounds>] <inFile>" in out.stdout
    out = subprocess.run([exe, str(tmp_path / "missing")], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 1 and "cannot open" in out.stderr
    bad = tmp_path / "bad"
    bad.write_text("AdjacencyGraph\n2\n1\n0\n" if name == "ST" else "EdgeArray\n0 1\n")
    out = subprocess.run([exe, str(bad)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "missing header" in out.stderr


def test_cpp_header_compiles_with_every_wrapper(tmp_path):
    """include/pbbs_bfs.hpp's wrappers (BFS, maximalIndependentSet, st) type-check
    against the C ABI; nothing runs."""
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "pbbs_bfs.hpp"\n'
        "int main(int argc, char** argv) {\n"
        "  if (argc < 3) return 0;\n"
        "  pbbs_amd::Context ctx(0);\n"
        "  pbbs_amd::HostGraph h = pbbs_amd::HostGraph::read(argv[1]);\n"
        "  pbbs_amd::Graph g(ctx, h);\n"
        "  std::vector<int> p = pbbs_amd::BFS(0, g);\n"
        "  std::vector<char> f = pbbs_amd::maximalIndependentSet(g);\n"
        "  pbbs_amd::EdgeArray e = pbbs_amd::EdgeArray::read(argv[2]);\n"
        "  std::vector<uint32_t> t = pbbs_amd::st(ctx, e);\n"
        "  std::vector<uint32_t> a = pbbs_amd::st(ctx, e, nullptr, PBBS_ST_ANY_FOREST);\n"
        "  return (int)(p.size() + f.size() + t.size() + a.size()) & 1;\n"
        "}\n")
    out = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                          "-I", os.path.join(ROOT, "include"), str(src)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
