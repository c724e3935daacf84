"""Benchmark of the MI355X rigid-body step.

Workload (the configuration the headline metric is quoted on): 4,096
independent worlds in one batch, body counts log-uniform in [16, 2048]
(seed 0xBB04), each a box pyramid (even) or 4x4 tower (odd) on its own ground
box, PGS_SOFT with 8 iterations, FP32, 63-bit Morton keys, hipGraph replay.
Beside it: the plain single world (1,000 cubes + ground, PGS_SOFT I=8, parity
mode) on the same GPU, and the CPU oracle on a bounded sample of the batch.

Prints one JSON line (see the contract in the repository's driver notes):
value = batch body-steps/s with all state resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "beat-box_amd"))

import beatbox  # noqa: E402
import scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes(stats: list, iters: int, kw: int) -> dict:
    """Algorithmic (compulsory) HBM bytes per step of each stage kernel, from
    the per-unit figures of the byte model: N bodies, A active, P pairs,
    M manifolds, C contacts, S = 84-byte compact state row, kw = key bytes."""
    N = sum(s["rigid_body_count"] for s in stats)
    A = sum(s["active_rigid_body_count"] for s in stats)
    P = sum(s["broad_phase_collision_count"] for s in stats)
    M = sum(s["collision_count"] for s in stats)
    Cc = sum(s["contact_count"] for s in stats)
    rounds = max((s["graph_color_count"] for s in stats), default=0)
    S = 84
    passes = 4 if kw == 4 else 8
    return {
        "counts": dict(N=N, A=A, P=P, M=M, C=Cc),
        "k_build": N * (12 + kw + 4) + passes * N * (3 * kw + 8) + N * (40 + 2 * kw)
        + (2 * N - 1) * 36 + N * 72 + 2 * N * S,
        "k_broad": N * 24 + P * 8,
        "k_narrow": P * 96 + M * 48 + Cc * 112,
        "k_post": M * 8 + A * 16 + rounds * M * 20 + N * 4,
        "k_pgs": (1 + 2 * iters) * (M * (36 + 2 * S) + Cc * 60) + A * 112,
        "k_avbd": 0,
        "k_finish": N * 16,
    }


def make_batch(worlds: int, seed: int, lo: int, hi: int):
    t0 = time.time()
    sizes = scenes.batch_world_sizes(seed, worlds, lo, hi)
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    b = beatbox.Batch()
    total = 0
    for w, n in enumerate(sizes):
        bodies = scenes.batch_world(seed, w, int(n))
        b.add_world(desc, bodies)
        total += len(bodies)
    log(f"batch: {worlds} worlds, {total} bodies, built in {time.time() - t0:.1f}s")
    return b, total


def time_batch(b, total: int, steps: int, warmup: int) -> float:
    b.step(warmup, use_graph=True)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps, use_graph=True)
    b.sync()
    return time.perf_counter() - t0


def plain_world(steps: int, warmup: int):
    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8))
    w = beatbox.Batch()
    bodies = scenes.cubes_grid()
    w.add_world(desc, bodies)
    w.step(warmup, use_graph=True)
    w.sync()
    t0 = time.perf_counter()
    w.step(steps, use_graph=True)
    w.sync()
    dt = time.perf_counter() - t0
    return len(bodies) * steps / dt, dt / steps * 1e3


def cpu_baseline(seed: int, budget_s: float):
    """The CPU oracle (single thread, -O2) on a bounded sample of the batch:
    the first worlds of the same seeded batch, stepped until the budget."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_bind import OracleWorld

    desc = beatbox.WorldDesc.make(**scenes.world_desc(iterations=8, morton_bits=63, quirks=0))
    sizes = scenes.batch_world_sizes(seed, 4096)
    ws, nb = [], 0
    for w in range(32):
        o = OracleWorld(desc)
        bodies = scenes.batch_world(seed, w, int(sizes[w]))
        o.add_bodies(bodies)
        ws.append(o)
        nb += len(bodies)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        for o in ws:
            o.step(1)
        steps += 1
    dt = time.perf_counter() - t0
    return dict(value=nb * steps / dt, unit="body-steps/s", cores=1, kind="port",
                sample=f"CPU oracle, 1 thread: worlds 0-31 of the seeded 4096-world batch "
                       f"({nb} bodies) x {steps} steps in {dt:.1f}s")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0xBB04)
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=6)
    ap.add_argument("--plain-steps", type=int, default=0, help="0: max(20, steps/2)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel on MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (worlds shard "
              "across GPUs without exchange; this build measures one GPU's share)", file=sys.stderr)
        return 2

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        ms = beatbox.copy_bandwidth(nbytes, 5)
        log(f"copy {args.copy_calibration} MiB: {ms:.3f} ms, {2 * nbytes / ms / 1e6:.1f} GB/s (read+write)")
        return 0

    copy_ms = beatbox.copy_bandwidth(1 << 30, 5)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"copy kernel: {copy_gbs:.0f} GB/s (practical HBM ceiling)")
    batch, total = make_batch(args.worlds, args.seed, 16, 2048)
    t = time_batch(batch, total, args.steps, args.warmup)
    log(f"device bytes: {batch.device_bytes / 2**30:.2f} GiB")
    value = total * args.steps / t
    ms_per_step = t / args.steps * 1e3
    log(f"batch: {ms_per_step:.3f} ms/step, {value / 1e6:.2f} M body-steps/s")

    # per-step latency and per-stage durations (hipEvents on the library stream)
    step_ms, kern_ms = batch.profile(args.profile_steps)
    p50 = float(np.percentile(step_ms, 50))
    p99 = float(np.percentile(step_ms, 99))
    stats = [batch.stats(w) for w in range(batch.world_count)]
    sb = stage_bytes(stats, 8, 8)
    per_launch_ms = {n: float(kern_ms[i]) / args.profile_steps for i, n in enumerate(beatbox.KERNEL_NAMES)}
    dominant = max(per_launch_ms, key=per_launch_ms.get)
    achieved = sb[dominant] / (per_launch_ms[dominant] * 1e-3) / 1e9
    stages = {n: dict(ms=round(per_launch_ms[n], 4),
                      gbs=round(sb[n] / (per_launch_ms[n] * 1e-3) / 1e9, 1) if per_launch_ms[n] > 0 else None)
              for n in beatbox.KERNEL_NAMES}
    log(f"stages: {stages}")

    plain_bs, plain_ms = plain_world(args.plain_steps or max(20, args.steps // 2), 5)
    log(f"plain: {plain_ms:.3f} ms/step, {plain_bs / 1e6:.3f} M body-steps/s, ratio {value / plain_bs:.1f}x")

    cpu = None if args.no_cpu else cpu_baseline(args.seed, args.cpu_budget)
    if cpu:
        log(f"cpu baseline: {cpu['value'] / 1e6:.3f} M body-steps/s")

    out = {
        "metric": "body-steps/s batched vs plain at fixed pose tol; p50/p99 step latency",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded box pyramids and towers)",
        "config": {
            "workload": "batched-worlds: 4096 worlds, bodies log-uniform [16,2048], PGS_SOFT I=8, hipGraph",
            "worlds": args.worlds,
            "bodies": total,
            "pairs": sb["counts"]["P"],
            "manifolds": sb["counts"]["M"],
            "contacts": sb["counts"]["C"],
            "plain_body_steps_per_s": plain_bs,
            "plain_ms_per_step": plain_ms,
            "batch_over_plain": value / plain_bs,
            "step_latency_ms_p50": p50,
            "step_latency_ms_p99": p99,
            "copy_kernel_gbs": copy_gbs,
            "stages": stages,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
