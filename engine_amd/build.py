"""Builds libmorphstore_hip.so in-tree with hipcc for gfx950.

Every HIP source is compiled separately (in parallel) and linked into one
shared library next to this file.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmorphstore_hip.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["ms_abi.hip", "ms_decode.hip", "ms_encode.hip", "ms_ops.hip", "ms_batch.hip", "ms_learn.hip", "ms_rel.hip", "ms_xfmt.hip", "ms_bwi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-I" + os.path.join(HERE, "..", "include")]


# The decode kernels branch on per-block bit widths held in scalar
# registers (wave-uniform): left unstructured, each such switch is a plain
# binary tree of scalar branches instead of a chain of exec-mask flow blocks.
UNIFORM_SWITCH = ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]
EXTRA = {"ms_decode.hip": UNIFORM_SWITCH, "ms_batch.hip": UNIFORM_SWITCH}


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
