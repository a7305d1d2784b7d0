"""Builds libnmc.so (the gfx950 HIP kernels + C-ABI) in-tree with hipcc.

Flags:
  --offload-arch=gfx950        MI355X (CDNA4) code objects only
  -ffp-contract=off            no fused multiply-add contraction: every float op rounds as written, the
                               precondition for bit-identical contact sets against a CPU build
  -target-feature -packed-fp32-ops
                               keep the compiler from forming packed-FP32 VALU instructions
"""

from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnmc.so")
SOURCES = ["collide.hip", "broad.hip", "mesh.hip", "match.hip", "xpbd.hip", "restitution.hip", "mlp.hip", "nmc_abi.hip", "noalloc.hip"]
HEADERS = ["nmc_math.h", "nmc_internal.h", "narrow.h", "sdf.h", "scan.h", "broad_common.h", "reduce.h", "xpbd_fast.h"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found; install ROCm or set HIPCC")


def flags() -> list[str]:
    return [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def build(force: bool = True, verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    if not force and os.path.exists(LIB):
        newest = max(os.path.getmtime(p) for p in srcs + [os.path.join(CSRC, h) for h in HEADERS])
        if os.path.getmtime(LIB) >= newest:
            return LIB
    # one hipcc per translation unit, in parallel, then one link
    common = [hipcc()] + [f for f in flags() if f != "-shared"]
    objs = [os.path.join(CSRC, os.path.splitext(s)[0] + ".o") for s in SOURCES]

    def compile_one(k: int) -> subprocess.CompletedProcess:
        return subprocess.run(common + ["-c", "-o", objs[k], srcs[k]], capture_output=True, text=True)

    jobs = max(1, min(len(srcs), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, range(len(srcs))))
    log = "".join(r.stdout + r.stderr for r in results)
    if any(r.returncode != 0 for r in results):
        raise RuntimeError("hipcc failed:\n" + log)
    tmp = LIB + ".tmp"
    r = subprocess.run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + r.stdout + r.stderr)
    for o in objs:
        os.remove(o)
    if verbose and (log or r.stdout or r.stderr):
        print(log + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True, verbose=True))
