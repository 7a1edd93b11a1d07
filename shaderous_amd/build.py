"""Build libshx.so (the HIP kernels and the C-ABI of include/shx.h) in-tree for gfx950, and the
N-API addon shaderous_amd/node/shx.node that binds it for Node.

``python -m shaderous_amd.build`` compiles every translation unit under shaderous_amd/csrc with
hipcc (in parallel), links shaderous_amd/libshx.so, then compiles the addon with gcc against the
system's Node headers (skipped, with a message, where they are absent). No GPU is needed to build.
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libshx.so")
NODE_DIR = os.path.join(HERE, "node")
ADDON = os.path.join(NODE_DIR, "shx.node")
NODE_INCLUDE = "/usr/include/node"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# the C ABI every translation unit includes
HEADER = os.path.join(os.path.dirname(HERE), "include", "shx.h")

DEVICE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # FMA only where one source expression forms a*b+c: every kernel that evaluates the same
    # expression rounds it the same way, so face fluxes agree bit for bit across kernels
    "-ffp-contract=on",
    # single-precision divide and sqrt by the hardware's ~1 ulp approximations
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


# per kernel, from hipcc's resource-usage remarks: registers, scratch, occupancy, LDS
_RESOURCE_KEYS = {"VGPRs": "vgprs", "AGPRs": "agprs", "ScratchSize [bytes/lane]": "scratch",
                  "Occupancy [waves/SIMD]": "occupancy", "VGPRs Spill": "vgpr_spill",
                  "LDS Size [bytes/block]": "lds"}


def _resources(stderr: str) -> list:
    out, cur = [], None
    for line in stderr.splitlines():
        m = re.search(r"remark:\s+(.*?)\s+\[-Rpass-analysis=kernel-resource-usage\]", line)
        if not m:
            continue
        text = m.group(1)
        if text.startswith("Function Name: "):
            cur = {"kernel": text[len("Function Name: "):]}
            out.append(cur)
        elif cur is not None and ": " in text:
            key, val = text.rsplit(": ", 1)
            if key in _RESOURCE_KEYS:
                cur[_RESOURCE_KEYS[key]] = int(val)
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src + ".o")
    res = os.path.join(BUILD, src + ".resources.json")
    deps = [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [HEADER]
    if (os.path.exists(obj) and os.path.exists(res)
            and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps)):
        return obj
    cmd = [HIPCC, *DEVICE_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")
    with open(res, "w") as f:
        json.dump(_resources(r.stderr), f)
    return obj


def kernel_resources() -> list:
    """Registers, scratch bytes per lane, occupancy and LDS of every kernel of the last build
    (hipcc's kernel-resource-usage remarks, recorded per translation unit)."""
    out = []
    for src in sources():
        path = os.path.join(BUILD, src + ".resources.json")
        if os.path.exists(path):
            with open(path) as f:
                out.extend(json.load(f))
    return out


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    if force:
        for f in os.listdir(BUILD):
            os.remove(os.path.join(BUILD, f))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, sources()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    build_addon()
    return LIB


def build_addon() -> str | None:
    """shx.node: gcc, the system Node headers, linked to libshx.so beside it (rpath $ORIGIN/..)."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        print(f"[build] no {NODE_INCLUDE}/node_api.h: the Node addon is not built", file=sys.stderr)
        return None
    src = os.path.join(NODE_DIR, "shx_napi.c")
    tmp = ADDON + ".tmp"
    cmd = ["gcc", "-shared", "-fPIC", "-O2", "-Wall", "-Wno-unused-parameter", f"-I{NODE_INCLUDE}",
           "-DNODE_GYP_MODULE_NAME=shx", src, "-o", tmp, f"-L{HERE}", "-lshx", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"addon build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, ADDON)
    return ADDON


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
