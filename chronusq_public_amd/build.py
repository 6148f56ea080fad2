"""Build libcqgpu.so (gfx950) in-tree.

Steps: generate the per-class ERI translation units (csrc/gen_eri.py -> csrc/_gen/),
compile every unit with hipcc in parallel, link one shared library next to this file.

    python -m chronusq_public_amd.build [--incremental] [--resource-usage] [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "_gen")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libcqgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str, obj: str, extra) -> str:
    cmd = [HIPCC] + HIP_FLAGS + list(extra) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    log = r.stdout + r.stderr
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{log}")
    return log


def build(incremental: bool = False, resource_usage: bool = False, jobs: int = 0) -> str:
    os.makedirs(BUILD, exist_ok=True)
    gen_script = os.path.join(CSRC, "gen_eri.py")
    subprocess.run([sys.executable, gen_script, GEN], check=True)
    srcs = sorted(glob.glob(os.path.join(GEN, "*.hip"))) + [os.path.join(CSRC, "gdirect.hip"), os.path.join(CSRC, "onee.hip"),
                                                              os.path.join(CSRC, "mlp.hip")]
    deps = [os.path.join(CSRC, f) for f in ("gdirect_device.h", "gdirect_kernels.h")] + [gen_script]
    dep_mtime = max(os.path.getmtime(d) for d in deps)
    header_mtime = os.path.getmtime(os.path.join(HERE, "..", "include", "cqgpu.h"))
    extra = ["-Rpass-analysis=kernel-resource-usage"] if resource_usage else []
    todo = []
    objs = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        objs.append(o)
        newest = max(os.path.getmtime(s), dep_mtime, header_mtime if s.endswith("gdirect.hip") else 0.0,
                     os.path.getmtime(os.path.join(CSRC, "onee.h")) if not s.startswith(GEN) else 0.0)
        if incremental and os.path.exists(o) and os.path.getmtime(o) > newest:
            continue
        todo.append((s, o))
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 4, 16)
    logs = {}
    # the biggest units first so the tail stays short
    todo.sort(key=lambda so: -os.path.getsize(so[0]))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, o, extra): s for s, o in todo}
        for f in cf.as_completed(futs):
            logs[futs[f]] = f.result()
    if resource_usage:
        with open(os.path.join(BUILD, "resource_usage.txt"), "w") as f:
            for s in sorted(logs):
                f.write(f"==== {os.path.basename(s)}\n{logs[s]}\n")
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--resource-usage", action="store_true")
    ap.add_argument("--jobs", type=int, default=0)
    a = ap.parse_args()
    print(build(a.incremental, a.resource_usage, a.jobs))
