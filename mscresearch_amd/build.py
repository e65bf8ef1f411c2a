"""Build recipe for the in-tree HIP library and the drop-in CLI.

Produces, inside the package directory:
  libmscmf.so   -- the C-ABI library (include/mscmf.h), kernels for gfx950
  bin/mscmf     -- `mscmf <input_dir> <output_dir> [static|dynamic]`
  bin/mscmf_scc -- the SCC contraction, query conversion/split and exact routing tools

Run as ``python -m mscresearch_amd.build`` or through ``__graft_entry__.build()``.
hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
ARCH = os.environ.get("MSCMF_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]
# Packed FP32 instructions are kept out of every device build.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["host_graph.cpp", "scc.cpp", "mf_kernels.hip", "mf_induced.hip", "mf_engine.cpp", "mf_tune.cpp", "mf_capi.cpp",
               "mf_noalloc.hip"]
CLI_SOURCES = ["mf_cli.cpp"]
SCC_CLI_SOURCES = ["mf_scc_cli.cpp"]

LIB = os.path.join(PKG, "libmscmf.so")
CLI = os.path.join(PKG, "bin", "mscmf")
SCC_CLI = os.path.join(PKG, "bin", "mscmf_scc")


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    if out.strip():
        sys.stderr.write(out + "\n")


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", path, "-o", obj]
    elif src in ("host_graph.cpp", "scc.cpp"):
        cmd = ["g++", *COMMON, "-pthread", "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(jobs: int | None = None) -> str:
    """Compile every source for gfx950 and link the library and CLI."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES + CLI_SOURCES + SCC_CLI_SOURCES))
    lib_objs = objs[: len(LIB_SOURCES)]
    cli_objs = objs[len(LIB_SOURCES): len(LIB_SOURCES) + len(CLI_SOURCES)]
    scc_objs = objs[len(LIB_SOURCES) + len(CLI_SOURCES):]
    tmp = LIB + ".tmp"
    _run([HIPCC, *DEVICE, "-shared", "-fPIC", "-o", tmp, *lib_objs, "-pthread"])
    os.replace(tmp, LIB)
    _run([HIPCC, "-o", CLI, *cli_objs, "-L" + PKG, "-lmscmf", "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    _run([HIPCC, "-o", SCC_CLI, *scc_objs, "-L" + PKG, "-lmscmf", "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    return LIB


if __name__ == "__main__":
    print(build())
