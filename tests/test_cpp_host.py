"""The C++ host mirror (include/bb_rigid_body_manager.hpp) compiled against
the C-ABI library and driven like the reference render loop drives its
RigidBodyManager."""
from __future__ import annotations

import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build(tmp_path):
    exe = str(tmp_path / "rbm_example")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "beat-box_amd", "examples", "rbm_example.cpp"),
                    "-L", os.path.join(REPO, "beat-box_amd"), "-lbeatbox_amd",
                    f"-Wl,-rpath,{os.path.join(REPO, 'beat-box_amd')}", "-o", exe], check=True)
    return exe


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_cpp_mirror_without_gpu_reports_no_device(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe, "2"], capture_output=True, text=True)
    assert r.returncode == 4 and "status 4" in r.stdout


@pytest.mark.gpu
def test_cpp_mirror_steps_on_gpu(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe, "60"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 61 and lines[-1].startswith("bodies 28")
    last = float(lines[-1].split()[-1])
    assert 0.3 < last < 0.6  # the bottom layer rests on the ground
