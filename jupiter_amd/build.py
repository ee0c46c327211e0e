"""Build libjgpm.so (HIP, gfx950) and the jupiter_gpm CLI in-tree with hipcc.

Usage: python -m jupiter_amd.build [--jobs N]
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libjgpm.so")
CLI = os.path.join(PKG, "jupiter_gpm")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    # packed fp32 instructions stay disabled in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
LIB_SOURCES = ["host_graph.cpp", "clique_kernels.hip", "sym_kernels.hip", "jgpm_api.hip", "dispatch.hip",
               "batch.hip", "serve.hip", "ingest.hip", "plan.cpp", "plan_kernels.hip", "noalloc.hip", "clique_huge.hip",
               # the reference-shaped C++ host API (Graph, VertexSet, Pattern, SubgraphSolver)
               "Graph.cc", "VertexSet.cc", "Pattern.cc", "solver.cc"]
CLI_SOURCES = ["main.cc"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    if src.endswith(".cpp") or src.endswith(".cc"):
        cmd = [HIPCC, "-x", "c++"] + [c for c in COMMON if not c.startswith("--offload-arch")] + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cli = [s for s in CLI_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs + cli))
    lib_objs = objs[: len(srcs)]
    cli_objs = objs[len(srcs):]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp] + lib_objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if cli_objs and len(cli_objs) == len(CLI_SOURCES):
        cmd = [HIPCC, "-o", CLI + ".tmp"] + cli_objs + ["-L" + PKG, "-ljgpm", "-Wl,-rpath,$ORIGIN", "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI link failed\n{r.stdout}\n{r.stderr}")
        os.replace(CLI + ".tmp", CLI)
    if verbose:
        print(f"built {LIB}" + (f" and {CLI}" if os.path.exists(CLI) else ""))
    return LIB


if __name__ == "__main__":
    j = None
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(j)
