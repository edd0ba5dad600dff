"""Build the gfx950 HIP library (libnanofiller.so) and the CLI in-tree.

Every translation unit is compiled by hipcc for gfx950 with
``-ffp-contract=off`` (each IEEE operation rounded once, as written) and
packed-FP32 instructions disabled; objects go to ``nanofiller_amd/_build``
and the shared library to ``nanofiller_amd/libnanofiller.so``.

Usage: ``python -m nanofiller_amd.build [--force] [-j N]``
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnanofiller.so")
CLI = os.path.join(HERE, "nanofiller")

SOURCES = [
    "nf_scan.hip",
    "nf_gen.hip",
    "nf_pair.hip",
    "nf_seg.hip",
    "nf_post.hip",
    "nf_net.hip",
    "nf_solve.hip",
    "nf_nodes.hip",
    "nf_api.cpp",
    "nf_contacts.cpp",
    "nf_network.cpp",
    "nf_io.cpp",
    "nf_state.cpp",
    "nf_conduct.cpp",
    "nf_nodes.cpp",
    "nf_noalloc.hip",
    "nf_reduce.hip",
]
HEADERS = ["nf_common.h", "nf_internal.h", "nf_host.h", "nf_seg.h"]
ARCH = os.environ.get("NF_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (need ROCm)")


def common_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-fPIC",
        "-std=c++17",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        "-Wno-unused-result",
        "-I", os.path.join(ROOT, "include"),
    ]


def _newer(target: str, deps: list[str]) -> bool:
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def _compile(src: str, force: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    deps = [path] + [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "nanofiller.h")]
    if force or _newer(obj, deps):
        cmd = [hipcc()] + common_flags() + ["-x", "hip", "-c", path, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        warn = "\n".join(l for l in r.stderr.splitlines()
                         if "packed-fp32-ops" not in l and l.strip())
        if warn and os.environ.get("NF_BUILD_VERBOSE"):
            print(warn, file=sys.stderr)
    return obj


def build(force: bool = False, jobs: int = 8) -> str:
    """Compile every HIP/C++ source for gfx950 and link libnanofiller.so."""
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    if force or _newer(LIB, objs):
        tmp = LIB + ".tmp"
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, LIB)
    # host build of the pair predicates for the CPU test suite (test only)
    harness = os.path.join(ROOT, "tests", "native", "band_harness.cpp")
    hlib = os.path.join(ROOT, "tests", "native", "libnf_band.so")
    if os.path.exists(harness) and (force or _newer(hlib, [harness, os.path.join(CSRC, "nf_common.h"),
                                                           os.path.join(CSRC, "nf_seg.h")])):
        cmd = [hipcc(), "--offload-host-only", "-x", "hip", "-O2", "-fPIC", "-shared", "-std=c++17",
               "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), harness, "-o", hlib]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"band harness build failed\n{' '.join(cmd)}\n{r.stderr}")
    cli_src = os.path.join(CSRC, "nf_cli.cpp")
    if os.path.exists(cli_src) and (force or _newer(CLI, [cli_src, LIB])):
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), cli_src,
               "-o", CLI, "-L", HERE, "-lnanofiller", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.j))


if __name__ == "__main__":
    main()
