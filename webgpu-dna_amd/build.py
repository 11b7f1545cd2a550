#!/usr/bin/env python3
"""Build libdna_amd.so (HIP, gfx950) in-tree.

hipcc flags:
  -ffp-contract=off                      no implicit FMA: FP32/FP64 results match the CPU
                                         restatement bit for bit (explicit fma kept)
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and sqrt
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib")
SOURCES = ["runtime.hip", "transport.hip", "canonical.hip", "irt.hip", "score.hip", "mlp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DNA_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-fhip-fp32-correctly-rounded-divide-sqrt",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def compile_one(src):
    obj = os.path.join(OUT, "obj", src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=True):
    os.makedirs(os.path.join(OUT, "obj"), exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    with ThreadPoolExecutor(max_workers=min(6, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    so = os.path.join(OUT, "libdna_amd.so")
    tmp = so + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, so)
    if verbose:
        print(f"built {so}")
    addon = build_napi(verbose)
    return so


NODE_INCLUDE = os.environ.get("DNA_NODE_INCLUDE", "/usr/include/node")


def build_napi(verbose=True):
    """Node-API addon lib/dna_amd.node over libdna_amd.so (skipped, with a message, when
    the Node headers are absent)."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print(f"node_api.h not found under {NODE_INCLUDE}: Node addon not built")
        return None
    out = os.path.join(OUT, "dna_amd.node")
    tmp = out + ".tmp"
    cmd = ["gcc", "-O2", "-Wall", "-shared", "-fPIC", f"-I{NODE_INCLUDE}", "-DNODE_GYP_MODULE_NAME=dna_amd",
           os.path.join(HERE, "napi", "dna_napi.c"), "-o", tmp, f"-L{OUT}", "-ldna_amd",
           "-Wl,-rpath,$ORIGIN", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Node addon build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
