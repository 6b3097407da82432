"""Build libblest.so (host C++ + HIP kernels for gfx950) and the `blest` CLI in-tree.

Usage: python -m blest_amd.build [--jobs N] [--debug]
Outputs: blest_amd/libblest.so, blest_amd/blest (objects under _build/).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build" / "blest"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = os.environ.get("BLEST_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["csc.cpp", "reorder.cpp", "bvss.cpp", "capi_host.cpp", "server.cpp", "autotune.cpp"]
DEVICE_SOURCES = ["device/dgraph.hip", "device/bfs.hip", "device/cc.hip", "device/mbfs.hip", "device/batch.hip",
                  "device/bfs_noalloc.hip", "device/bvss_build.hip"]
CLI_SOURCE = "cli/blest_main.cpp"

COMMON = ["-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    # gfx950: keep the compiler from forming packed-FP32 VALU ops.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")


def _newer(src: Path, obj: Path, deps: list[Path]) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return src.stat().st_mtime > t or any(d.stat().st_mtime > t for d in deps)


def build(jobs: int = 0, debug: bool = False, force: bool = True) -> Path:
    opt = ["-O0", "-g"] if debug else ["-O3"]
    BUILD.mkdir(parents=True, exist_ok=True)
    (BUILD / "device").mkdir(exist_ok=True)
    (BUILD / "cli").mkdir(exist_ok=True)
    headers = list(CSRC.rglob("*.hpp")) + [ROOT / "include" / "blest.h"]
    tasks = []
    for s in HOST_SOURCES:
        src, obj = CSRC / s, BUILD / (s + ".o")
        if force or _newer(src, obj, headers):
            tasks.append(["g++", *COMMON, *opt, "-pthread", "-c", str(src), "-o", str(obj)])
    for s in DEVICE_SOURCES:
        src, obj = CSRC / s, BUILD / (s + ".o")
        if force or _newer(src, obj, headers):
            tasks.append([HIPCC, *COMMON, *opt, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, tasks))
    objs = [str(BUILD / (s + ".o")) for s in HOST_SOURCES + DEVICE_SOURCES]
    lib = PKG / "libblest.so"
    tmp = PKG / "libblest.so.tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp), "-pthread",
          "-Wl,-soname,libblest.so"])
    os.replace(tmp, lib)
    # measurement probe of level-synchronisation costs (no library code)
    probe = CSRC / "probes" / "level_sync_probe.hip"
    if probe.exists():
        _run([HIPCC, "-std=c++20", *opt, *HIP_FLAGS, str(probe), "-o", str(PKG / "level_sync_probe")])
    cli_src = CSRC / CLI_SOURCE
    if cli_src.exists():
        exe = PKG / "blest"
        _run(["g++", *COMMON, *opt, str(cli_src), "-o", str(exe), f"-L{PKG}", "-lblest",
              "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM / 'lib'}", "-pthread"])
    return lib


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--debug", action="store_true")
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.debug, force=not a.incremental))


if __name__ == "__main__":
    main()
