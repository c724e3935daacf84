"""bench.py's host logic, without a GPU: the roofline's measured traffic is
reported only for a run with the committed counter profile's step window and
launch parameters, and more than one GPU is refused with a message."""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def _profile():
    with open(bench.TRAFFIC_FILE) as f:
        return json.load(f)


def test_traffic_only_for_the_profiled_window_and_parameters():
    meta = _profile()
    params = dict(meta["launch_params"])
    win = meta["window"]
    model = meta["model_bytes_per_launch"]["k_pgs"]
    n_wide = next(int(k.split("@")[1]) for k in meta["kernels"]
                  if k.startswith(f"bb::k_pgs<{params['lane_block']}>@"))
    args = argparse.Namespace(steps=win["steps"], warmup=win["warmup"], worlds=4096)
    out = bench.traffic_fields("k_pgs", args, model, params, n_wide)
    assert out["traffic"] and 0.1 < out["traffic_over_model"] < 2.0, out
    # per lane: the kernel's read + write bytes of the two launches of the step
    want = sum(meta["kernels"][k]["read_bytes"] + meta["kernels"][k]["write_bytes"]
               for k in (f"bb::k_pgs<{params['pgs_block']}>@{4096 - n_wide}",
                         f"bb::k_pgs<{params['lane_block']}>@{n_wide}"))
    assert abs(out["traffic"] - want) < 1.0
    # another window or other launch parameters: no traffic, the profiled figures beside it
    other = dict(params, pgs_lds_rows=params["pgs_lds_rows"] + 128)
    for a, p in ((argparse.Namespace(steps=240, warmup=10, worlds=4096), params), (args, other)):
        out = bench.traffic_fields("k_pgs", a, model, p, n_wide)
        assert out["traffic"] is None and out["traffic_over_model"] is None
        assert out["traffic_profiled"]["launch_params"] == params


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0 and "one GPU" in r.stderr, (r.returncode, r.stderr[-500:])
