"""bench.py and the autotuner's command lines, CPU only (nothing here reaches
a GPU call)."""
import os
import subprocess
import sys

from conftest import ROOT


def run(*args):
    return subprocess.run([sys.executable, *args], cwd=ROOT, capture_output=True, text=True,
                          timeout=120)


def test_bench_refuses_more_than_one_gpu():
    r = run(os.path.join(ROOT, "bench.py"), "--gpus", "2")
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr and r.stdout == ""


def test_bench_help_names_every_mode():
    r = run(os.path.join(ROOT, "bench.py"), "--help")
    assert r.returncode == 0
    for word in ("single", "batch", "serve", "ingest", "mis", "st", "--no-mis", "--no-st",
                 "--spread", "--workload"):
        assert word in r.stdout


def test_autotune_cli_needs_a_graph_and_a_table():
    r = run("-m", "pbbsbench_amd.autotune", "--out", "t.json")
    assert r.returncode == 2 and "one of the arguments" in r.stderr
    r = run("-m", "pbbsbench_amd.autotune", "--help")
    assert r.returncode == 0
    for word in ("--rmat", "--grid", "--graph", "--batch", "--wide-offsets"):
        assert word in r.stdout
