"""Build recipe for libbeatbox_amd.so (gfx950).

hipcc flags: -ffp-contract=off keeps every float operation in source order
(no fused multiply-add), which the bit-exact comparison with the CPU oracle
relies on; packed-fp32 instructions are disabled for all device code.

The units compile in parallel into _build/ and link into the library:
bb_step.hip (team kernel, launch sequence), bb_step_block.hip once per
per-world workgroup size (BB_BLOCK), bb_capi.cpp, bb_server.cpp and
bb_noalloc.hip.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libbeatbox_amd.so")
OBJ_DIR = os.path.join(HERE, "_build")
BLOCKS = (128, 256, 512, 1024)
# (source, extra flags, object name)
UNITS = [("bb_step.hip", [], "bb_step.o")] + [
    ("bb_step_block.hip", [f"-DBB_BLOCK={b}"], f"bb_step_block_{b}.o") for b in BLOCKS] + [
    ("bb_capi.cpp", [], "bb_capi.o"), ("bb_server.cpp", [], "bb_server.o"), ("bb_noalloc.hip", [], "bb_noalloc.o")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)

    def compile_one(unit):
        src, extra, obj = unit
        out = os.path.join(OBJ_DIR, obj)
        cmd = [HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return out

    jobs = max(1, min(len(UNITS), os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, UNITS))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
