"""Builds the CPU emulation of the step (tests/cpu_emu/_build/emu_main).

The kernels of beat-box_amd/csrc/bb_step.hip and the host code of the C-ABI
(bb_capi.cpp, bb_server.cpp) are compiled unchanged by g++, against the HIP
emulation header in tests/cpu_emu/hip/, with AddressSanitizer.  emu_main
(main.cpp) steps the worlds of a scene file through the C-ABI.
"""
from __future__ import annotations

import fcntl
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(REPO, "beat-box_amd", "csrc")
OUT_DIR = os.path.join(HERE, "_build")
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address",
         "-ffp-contract=off", "-w", "-I", HERE]
# (source, compile as C++, extra flags, object name): bb_step_block.hip once
# per workgroup size, as the library's build does
UNITS = [(os.path.join(CSRC, "bb_step.hip"), True, [], "bb_step.o")] + [
    (os.path.join(CSRC, "bb_step_block.hip"), True, [f"-DBB_BLOCK={b}"], f"bb_step_block_{b}.o")
    for b in (128, 256, 512, 1024)] + [
    (os.path.join(CSRC, "bb_capi.cpp"), False, [], "bb_capi.o"),
    (os.path.join(CSRC, "bb_server.cpp"), False, [], "bb_server.o"),
    (os.path.join(HERE, "emu.cpp"), False, [], "emu.o"), (os.path.join(HERE, "main.cpp"), False, [], "main.o")]


def build() -> str:
    """Builds emu_main; test processes running side by side (pytest -n) take
    turns through a lock file, so no two write the same objects at once."""
    os.makedirs(OUT_DIR, exist_ok=True)
    with open(os.path.join(OUT_DIR, ".lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        return _build()


def _build() -> str:
    exe = os.path.join(OUT_DIR, "emu_main")

    def compile_one(unit):
        src, as_cxx, extra, name = unit
        obj = os.path.join(OUT_DIR, name)
        cmd = ["g++", *FLAGS, *extra, "-c"] + (["-x", "c++"] if as_cxx else []) + [src, "-o", obj]
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, UNITS))
    subprocess.run(["g++", *FLAGS, *objs, "-o", exe, "-lpthread"], check=True)
    return exe


if __name__ == "__main__":
    print(build())
