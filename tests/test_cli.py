"""The `blest` CLI (reference main.cu / Benchmark.cuh flow over the C ABI)."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "blest_amd" / "blest"


def run(*args, cwd=None):
    return subprocess.run([str(EXE), *args], capture_output=True, text=True, cwd=cwd, timeout=120)


@pytest.mark.parametrize("args,msg", [
    (["-g", "x", "-j", "1", "-w", "8"], "Missing required option: -d"),
    (["-d", "/tmp", "-j", "1", "-w", "8"], "Missing required option: -g"),
    (["-d", "/tmp", "-g", "x", "-w", "8"], "Missing required option: -j"),
    (["-d", "/tmp", "-g", "x", "-j", "1"], "Missing required option: -w"),
    (["-d", "/tmp", "-g", "x", "-j", "2", "-w", "8"], "Invalid value for -j"),
    (["-d", "/tmp", "-g", "x", "-j", "1", "-w", "abc"], "Invalid value for -w"),
    (["-d", "/tmp", "-g", "x", "-j", "1", "-w", "8", "--bogus"], "Unknown argument"),
])
def test_cli_argument_errors(args, msg):
    r = run(*args)
    assert r.returncode == 1
    assert msg in r.stderr and "Usage:" in r.stderr


def test_cli_missing_matrix(tmp_path):
    r = run("-d", str(tmp_path), "-g", "absent", "-j", "1", "-w", "64", "--no-cache")
    assert r.returncode == 1 and "Failed to open" in r.stderr


def test_cli_help():
    r = run("-h")
    assert r.returncode == 0 and "Usage:" in r.stderr


@pytest.mark.gpu
def test_cli_bfs_end_to_end(tmp_path, blest, oracle):
    # a small undirected MatrixMarket graph and a sources file, as the reference expects
    g = blest.CSC.rmat(10, 8, seed=3)
    cp, rows = g.arrays()
    lines = ["%%MatrixMarket matrix coordinate pattern symmetric", f"{g.n} {g.n} {len(rows) // 2}"]
    for j in range(g.n):
        for i in rows[cp[j]:cp[j + 1]]:
            if i > j:
                lines.append(f"{j + 1} {i + 1}")
    (tmp_path / "tiny.mtx").write_text("\n".join(lines) + "\n")
    srcs = g.pick_roots(4, seed=1)
    (tmp_path / "tiny.txt").write_text("\n".join(str(s) for s in srcs) + "\n")
    r = run("-d", str(tmp_path), "-g", "tiny", "-j", "1", "-w", "256")
    assert r.returncode == 0, r.stderr
    got = re.findall(r"Source: (\d+) - Number of levels processed: (\d+) - Total visited: (\d+)", r.stdout)
    assert len(got) == 4
    for s, lev, vis in got:
        want = oracle.bfs(g.n, cp, rows, int(s))
        assert int(lev) == want.max() + 1 and int(vis) == int((want >= 0).sum())
    assert (tmp_path / "tiny_csc.bin").exists() and (tmp_path / "tiny_ordering.bin").exists()
    csv = (tmp_path / "tiny.csv").read_text().splitlines()
    assert csv[0].startswith("N\tNNZ") and csv[2].startswith("Source\tTotalLevels")
    # second run reuses the caches
    r2 = run("-d", str(tmp_path), "-g", "tiny", "-j", "1", "-w", "256", "--kernel", "CC")
    assert r2.returncode == 0 and "CSC read from binary." in r2.stdout and "Ordering read from binary." in r2.stdout
    comps = int(re.search(r"Number of components identified: (\d+)", r2.stdout).group(1))
    assert comps == oracle.cc(g.n, cp, rows)[1]
